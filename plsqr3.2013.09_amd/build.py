"""Build recipe for libplsqr_hip.so (gfx950).

    python plsqr3.2013.09_amd/build.py
Compiles in-tree so the built files travel with the source tree.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libplsqr_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PLSQR_ARCH", "gfx950")

# -ffp-contract=off: products and sums are rounded separately, as in the
# reference's x86-64 build; explicit fma() calls are unaffected.
# -packed-fp32-ops off: no v_pk_*_f32 generation.
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-I", INCLUDE, "-I", CSRC]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


SOURCES = (("plsqr_api.hip", DEVICE), ("plsqr_batch.hip", DEVICE), ("plsqr_cnn.hip", DEVICE),
           ("plsqr_relayout.hip", DEVICE), ("tomo_io.cpp", []),
           ("tile_build.cpp", []), ("batch_build.cpp", []))


def build_lib(out=LIB):
    tmp = out + ".tmp"
    obj_dir = os.path.join(HERE, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    jobs = []
    for src, extra in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(obj_dir, src + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else ["-x", "c++"]
        jobs.append(([HIPCC] + COMMON + extra + lang + ["-c", path, "-o", obj], obj))
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", "8"))))
    with ThreadPoolExecutor(workers) as ex:
        list(ex.map(lambda j: run(j[0]), jobs))
    objs = [o for _, o in jobs]
    run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, out)
    return out


def main():
    build_lib()


if __name__ == "__main__":
    sys.exit(main())
