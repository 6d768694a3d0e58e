// Host-side layout of a batch of independent systems (internal to
// libplsqr_hip.so).
//
// The batch is one block-diagonal operator: problem p owns the global rows
// [row0[p], row0[p] + m_k + m_d) (kernel rows, then damping rows) and the
// global columns [col0[p], col0[p] + n).  Inside a problem the columns are
// renumbered into bricks of voxels (brick_order) exactly as for one system,
// and every device vector holds K right-hand sides interleaved, element
// (i, k) at i*K + k, so that one gather fetches all K values of an index.
//
// Work is cut into pieces that never straddle two problems, so that every
// partial sum belongs to exactly one problem:
//   row chunks    (mode 1): runs of kernel rows or of damping rows;
//   tiles         (mode 2): one brick each, with its ray segments and slots;
//   column chunks (x/w update): runs of columns.
// The (problem, right-hand side) state reduces the partials of its problem's
// pieces in a fixed order, so results do not depend on scheduling.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "plsqr_hip.h"
#include "tile_build.h"

namespace plsqr {

// One problem in local numbering (rows 0..m_k+m_d, new columns 0..n).
struct ProblemLayout {
    int m_k = 0, m_d = 0, n = 0;
    std::vector<int> inv;         // inv[new column] = old column
    std::vector<int> rperm;       // rperm[new kernel row] = old kernel row (rays sorted by tiles)
    std::vector<int> rp, ci;      // CSR of [A_k ; D], new columns ascending in kernel rows
    std::vector<double> rv;
    std::vector<int> dcp, dri;    // damping CSC in new column order, damping-local rows
    std::vector<double> dcv;
    TileLayout tiles;             // S row slices, items t*S + s
    int S = 1;
};

struct BatchPlan {
    int P = 0, TW = 128;
    std::vector<ProblemLayout> prob;
    // offsets (P + 1 entries each)
    std::vector<long long> row0, col0, nnz0, dnnz0, slot0, seg0, cut0;
    // mode-1 segment groups and entries (RowTiles of each problem's tiles)
    std::vector<long long> grp0, ent0;
    std::vector<long long> pofs;         // start of each problem's slice partials (in columns)
    std::vector<int> item0, rch0, cch0;  // items = (tile, slice) pairs
    std::vector<int> nslice;             // row slices per problem
    // mode-2 launch order: workgroup b runs item blk_item[b] (-1: none).
    // Workgroup b lands on XCD b % 8; every row slice is given to one XCD
    // and its items are queued there together, so the slice's rows of u
    // stay in that XCD's L2 while its items run.
    std::vector<int> blk_item;
    // mode-1 per-ray sums: workgroup b runs row chunk blk_rch[b] (-1: none),
    // a slice's kernel-row chunks on the XCD that ran the slice's items
    std::vector<int> blk_rch;
    // pieces, global numbering
    std::vector<int> rch_row, rch_prob;  // row chunks: rch_row has count + 1 entries
    std::vector<uint8_t> rch_kind;       // 0 kernel rows, 1 damping rows
    std::vector<int> cch_col, cch_prob;  // column chunks: cch_col has count + 1 entries
    long long rows = 0, cols = 0, nnz = 0, dnnz = 0, slots = 0, segs = 0, cuts = 0, partials = 0;
    long long groups = 0, ents = 0;
    int items = 0;
    int max_seg_per_item = 0;
};

// Checks a system's arrays on the host: pointer arrays monotone and ending at
// nnz, indices in range, values finite.  Returns a PLSQR_* code with *err set.
int validate_system(const plsqr_system *s, std::string *err);

// Builds the plan.  grids[p] may describe problem p's voxel grid (brick tiles
// of 4 x 4 x 8) or be zero-sized / absent (tiles of TW consecutive columns).
// Kernel rows are cut into slices of about slice_rows rows (balanced by
// nonzeros); xcds is the number of L2 domains to queue the slices on.
int build_batch_plan(const plsqr_system *const *systems, const plsqr_grid *grids, int P, int TW,
                     int row_chunk, int damp_chunk, int col_chunk, int slice_rows, int xcds, BatchPlan *plan,
                     std::string *err);

}  // namespace plsqr
