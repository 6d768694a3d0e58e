// Host-side construction of the tiled mode-2 layout (internal to libplsqr_hip.so).
//
// Columns are renumbered so that a tile of consecutive new column indices is
// a compact brick of voxels (4 x 8 x 8 when the grid is known; the batched
// operator uses 4 x 4 x 8).  For every
// (row slice s, tile t) "item" the kernel gathers u once per ray that crosses
// the tile (a segment) and reuses it for every voxel of the tile the ray
// touches.
//
// Per item: the segments' rows (seg_ray), and the tile's nonzeros in "slots":
// slot (j, v) = slot_base[item] + j * TW + v holds the j-th nonzero (rows
// ascending) of the tile's voxel v within slice s, as (value, local segment
// index).  vox_len[item * TW + v] nonzeros per voxel; slots past vox_len are
// padding that the kernel never reads.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <string>
#include <thread>
#include <vector>

#include "plsqr_hip.h"

namespace plsqr {

// Runs f(b, e) over a split of [0, n) on up to PLSQR_HOST_THREADS (default:
// all hardware threads, at most 64) host threads.
template <class F>
void par_for(long long n, F &&f) {
    unsigned nt = std::thread::hardware_concurrency();
    if (const char *e = std::getenv("PLSQR_HOST_THREADS")) nt = (unsigned)std::max(1, atoi(e));
    nt = std::max(1u, std::min(nt, 64u));
    if (n < 64) nt = 1;
    std::vector<std::thread> th;
    const long long chunk = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        const long long b = (long long)t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([&f, b, e] { f(b, e); });
    }
    for (auto &x : th) x.join();
}

struct TileLayout {
    int S = 1;   // row slices
    int T = 0;   // tiles
    int TW = 128;
    std::vector<int> perm;        // perm[old column] = new column
    std::vector<int> inv;         // inv[new column] = old column
    std::vector<int> tile_start;  // T+1, in new column numbering
    std::vector<int> slice_bounds;  // S+1 kernel rows
    std::vector<long long> slot_base;  // items+1 (item = t*S + s)
    std::vector<long long> seg_start;  // items+1
    std::vector<uint16_t> vox_len;     // items * TW
    std::vector<int> seg_ray;          // segments
    std::vector<double> sval;          // slots
    std::vector<uint16_t> sseg;        // slots
    // Round cuts for items with more than CUT segments: the kernel stages the
    // segments CUT (or a multiple) at a time, and vox_cut[(cut_base[item] +
    // q) * TW + v] = the number of voxel v's slots whose segment is below
    // (q + 1) * CUT, so every round is a plain loop over a known slot range.
    static constexpr int CUT = 1024;
    std::vector<long long> cut_base;   // items+1
    std::vector<uint16_t> vox_cut;     // cut_base[items] * TW
    long long n_seg = 0, n_slot = 0;
    int max_seg_per_item = 0;
    // Slots by position (compact_tile_layout): inside every item the voxels
    // are taken by descending vox_len ("positions"; pos_vox[item * TW + p] =
    // the voxel at position p), so the positions still active in row j are
    // a prefix of the row.  Rows are either kept TW wide (slot (j, p) =
    // slot_base[item] + j * TW + p) or packed: row j then holds only its
    // active positions, from row_off[row_base[item] + j] (relative to
    // slot_base[item]), and the slots carry no padding.  vox_len and vox_cut
    // are indexed by position in both.
    bool compact = false;
    bool packed = false;
    std::vector<uint8_t> pos_vox;      // items * TW
    std::vector<long long> row_base;   // items+1 (packed rows)
    std::vector<int> row_off;          // row_base[items] (packed rows)
};

// Rewrites a layout's slots by position, with packed rows if `pack` (no-op
// if already done).  In the default form each voxel's j-th nonzero sits at
// a fixed stride, so a row of 64 voxels of which only a few are still
// active costs whole cache lines; by position the active lanes of a row are
// a prefix, and packed rows make each item one contiguous stream.  The
// segments, the sums and their order are unchanged: every voxel still adds
// its slots in ascending j.
void compact_tile_layout(TileLayout *L, bool pack = true);

// Mode-1 layout on the same (tile, row slice) items.  Inside an item every
// segment (one ray's entries in the tile) is a list of (voxel, value) pairs
// in ascending voxel order; the segments keep the item's ray order and are
// dealt to groups of 64, one per lane (a wavefront), with each segment's
// entries stored together: entry k of lane l sits at ent_base[group] + (the
// lengths of lanes < l) + k.  No padding is stored.  The kernel sums each
// segment against the tile of x staged in LDS and writes the partial to
// segv[seg_start[item] + segment index]; the combine pass adds a ray's
// partials in tile order through ray_ptr/ray_seg.  Keeping ray order makes
// neighbouring rays' partials neighbours in segv, so the combine's reads
// share cache lines.
struct RowTiles {
    static constexpr int G = 64;       // segments per group (one wavefront)
    std::vector<int> grp_base;         // items+1: first group of each item
    std::vector<long long> ent_base;   // groups+1: first entry of each group
    std::vector<uint8_t> seg_len;      // groups*G: entries of each lane's segment (0: no segment)
    // Tiles of more than 256 voxels: a segment's entries ascend in voxel, so
    // the entries below voxel 256, 512, 768 are prefixes of it, and seg_lo
    // holds their lengths, cut_bytes per segment: one byte (tiles of at most
    // 512 voxels: the voxel of entry k is ent_vox + 256 from k = seg_lo on),
    // or a 32-bit word (at most 1024 voxels) whose byte j is the number of
    // entries below voxel 256 (j + 1) (entry k is ent_vox + 256 times the
    // number of bytes <= k; bytes past the tile's width hold 255).  Empty,
    // cut_bytes 0, for tiles of at most 256 voxels.
    int cut_bytes = 0;
    std::vector<uint8_t> seg_lo;       // groups*G*cut_bytes
    // Segments sorted by length inside windows (build_row_tiles): the
    // partial of lane l of group g goes to segv[seg_start[item] +
    // seg_dst[g*G + l]], the segment's index in the item's ray order, so the
    // partials keep ray order whatever the lanes' order.  Empty: lane order
    // is ray order, and the partial goes to (g - grp_base[item])*G + l.
    std::vector<uint16_t> seg_dst;     // groups*G
    std::vector<uint8_t> ent_vox;      // entries: voxel inside the tile (mod 256)
    std::vector<double> ent_val;       // entries
    std::vector<int> ray_ptr;          // m_k+1
    std::vector<int> ray_seg;          // n_seg: segment indices, tile order per ray
    long long n_ent = 0;
    int n_grp = 0;
};

// Builds the mode-1 layout from a mode-2 layout (same items and segments).
// Inside an item the segments are taken in windows of `window` (ray order
// between windows), sorted by descending length inside each window so that a
// wavefront's 64 segments have similar lengths (window <= 1: ray order).
// With sorted windows the partials still keep ray order (seg_dst), so each
// segment's sum and each ray's sum of partials are the same as in ray order.
// Tiles may hold up to 1024 voxels (seg_lo above).  Returns false with *err
// set when a segment holds more than 255 entries, the group count overflows,
// a tile is wider than 1024 or the layout is already compacted (build the
// mode-1 layout first, then compact_tile_layout).
bool build_row_tiles(const TileLayout &L, int m_k, RowTiles *out, std::string *err, int window = 1);

// Host evaluation of y = A_k x (kernel rows, new column numbering) through
// the mode-1 layout, adding every ray's partials in tile order; the check
// of the layout against the CSC before upload.
void row_tiles_apply(const TileLayout &L, const RowTiles &R, int m_k, const double *x, double *y);

// Reorders the columns inside every tile by descending kernel nonzero count
// (stable), updating perm (old -> new) in place.  The tiled mode-2 kernel
// reads a voxel-interleaved row of slots per step; with long columns first
// the lanes still active in a late step are contiguous, so the cache lines
// it fetches are full instead of mostly padding.
void order_tile_columns(const plsqr_system *s, const std::vector<int> &tile_start, std::vector<int> &perm);

// Kernel rows renumbered so that rays crossing the same bricks are
// neighbours (a stable sort, equal keys keep the loaded order): with the
// voxel grid known, by a 6-D Morton code of the ray's two end voxels (found
// from its columns: the extremes along its longest axis); without one, by
// the first and last tile its columns fall in.  Neighbouring rays then share
// cache lines of u and of the segment partials inside a brick.  With slice bounds (row ranges of the loaded order) the
// sort stays inside each slice, so every slice keeps its rows and its mix of
// ray directions.  sys gets the renumbered kernel CSC (rows ascending in
// every column); the damping arrays and b are the caller's.  rperm[new] =
// old kernel row, rinv[old] = new.
// The system with its kernel rows sorted (rperm/rinv over the first m_k
// entries) and its damping rows sorted by the new position of the cell that
// owns them (the row's first column), so that the damping rows a brick of
// columns touches are neighbours in u (entries m_k.. of rperm/rinv, offset by
// m_k).  rperm[new row] = old row, rinv[old row] = new row, over all m rows.
struct RaySorted {
    plsqr_system sys;
    std::vector<int> ri, rperm, rinv;
    std::vector<double> kv;
    std::vector<int> drp, dci, dcr;    // damping CSR (row_ptr, col_idx) and CSC row_idx in the new row order
    std::vector<double> dv, dcv;       // their values
};

// sort_damping = false leaves the damping rows in place (rperm/rinv then
// cover the m_k kernel rows only).
void sort_rays(const plsqr_system *s, const std::vector<int> &perm, const std::vector<int> &tile_start,
               const std::vector<int> &slice_bounds, const plsqr_grid *grid, RaySorted *out,
               bool sort_damping = true);

// Brick permutation for an nx*ny*nz grid (x fastest): bricks of bx*by*bz
// voxels, voxels inside a brick x fastest.  Brick order: 1 = z fastest
// (vertical stacks of bricks, then x, then y), 2 = Morton (Z-order) of the
// brick coordinates, 0 = x fastest.  Workgroups run roughly in brick order,
// so the order sets how far apart in time one ray's bricks are processed.
// Fills perm (old -> new) and the tile starts (one tile per brick).
void brick_order(int nx, int ny, int nz, int bx, int by, int bz, int order,
                 std::vector<int> &perm, std::vector<int> &tile_start);

// Brick order 1 with the bricks grouped into super-bricks of qx * qy * qz
// bricks: super-bricks z fastest, then x, then y, and inside each its bricks
// z fastest, then x, then y (qx = qy = qz = 1 is order 1 itself, and
// 1 x 1 x 2 numbers the columns as order 1 does).  Every super-brick is a
// run of consecutive bricks, so its columns are a contiguous range: tiles of
// mode 1 that are unions of mode 2's brick tiles under one column order.
// Fills perm, the brick starts (tile_start) and the super-brick starts
// (super_start, a subset of tile_start).
void brick_order_grouped(int nx, int ny, int nz, int bx, int by, int bz, int qx, int qy, int qz,
                         std::vector<int> &perm, std::vector<int> &tile_start, std::vector<int> &super_start);

// Builds the layout from the kernel CSC of `s`.  perm/tile_start describe the
// column renumbering and the tiles (tile widths <= TW); slice_bounds the row
// slices.  Returns false with *err set when an item needs more than 65535
// segments (the local segment index is 16-bit) or a tile is wider than TW.
bool build_tile_layout(const plsqr_system *s, const std::vector<int> &perm,
                       const std::vector<int> &tile_start, const std::vector<int> &slice_bounds,
                       int TW, TileLayout *out, std::string *err);

// Host evaluation of the per-slice column sums P[s*n + new_col] from the
// layout, used to check the layout against the CSC before it is uploaded.
// The slots are visited round by round (CUT segments per round) through the
// round cuts, as the batch kernel does; returns false if a cut disagrees with
// the segments of the slots it bounds.  Reads either slot form.
bool tile_layout_apply(const TileLayout &L, int n, const double *u, double *P);

// Compares tile_layout_apply with the CSC of s on a fixed test vector (the
// layout's own permutation, slices and cuts); false with *err on a mismatch.
bool tile_layout_self_check(const plsqr_system *s, const TileLayout &L, std::string *err);

// Compares row_tiles_apply with the CSC of s on a fixed test vector; false
// with *err on a mismatch.
bool row_tiles_self_check(const plsqr_system *s, const TileLayout &L, const RowTiles &R, std::string *err);

}  // namespace plsqr
