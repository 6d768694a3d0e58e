// The thread-local message behind plsqr_last_global_error, for calls that
// have no handle (internal to libplsqr_hip.so).
#pragma once

namespace plsqr {

// Records msg as the last global error and returns code.
int global_fail(int code, const char *msg);

}  // namespace plsqr
