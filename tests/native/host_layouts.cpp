// Host-side code of libplsqr_hip.so under AddressSanitizer and
// UndefinedBehaviorSanitizer: the generator, the reference file formats, the
// damping tool, the ASCII kernel reader, the tiled layouts (8- and 16-bit
// index packing, round cuts) and the batch plan, each run on small systems
// and checked against a plain recomputation.  Built with g++ from
// tomo_io.cpp, tile_build.cpp and batch_build.cpp alone (no HIP), by
// tests/test_host_sanitize.py.  Usage: host_layouts <scratch dir>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "batch_build.h"
#include "plsqr_hip.h"
#include "tile_build.h"

static int failures = 0;

#define CHECK(cond, ...)                                   \
    do {                                                   \
        if (!(cond)) {                                     \
            std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
            std::fprintf(stderr, __VA_ARGS__);             \
            std::fprintf(stderr, "\n");                    \
            ++failures;                                    \
        }                                                  \
    } while (0)

static plsqr_gen_params params(int nx, int ny, int nz, long long rays, int geom, unsigned long long seed) {
    plsqr_gen_params p;
    std::memset(&p, 0, sizeof p);
    p.nx = nx;
    p.ny = ny;
    p.nz = nz;
    p.n_rays = rays;
    p.geometry = geom;
    p.seed = seed;
    p.sigma = 1e-3;
    p.lambda = 0.5;
    p.w_xx = p.w_yy = p.w_zz = 0.25;
    p.n_blobs = 5;
    p.blob_width = 2.0;
    return p;
}

// y = A_k x from the kernel CSC
static std::vector<double> csc_apply(const plsqr_system *s, const std::vector<double> &x) {
    std::vector<double> y((size_t)s->m_k, 0.0);
    for (int c = 0; c < s->n; ++c)
        for (int e = s->k_col_ptr[c]; e < s->k_col_ptr[c + 1]; ++e) y[(size_t)s->k_row_idx[e]] += s->k_val[e] * x[(size_t)c];
    return y;
}

static std::vector<int> slice_bounds(const plsqr_system *s, int nsl) {
    std::vector<long long> cnt((size_t)s->m_k + 1, 0);
    for (long long e = 0; e < s->nnz_k; ++e) cnt[(size_t)s->k_row_idx[e] + 1]++;
    for (int r = 0; r < s->m_k; ++r) cnt[(size_t)r + 1] += cnt[(size_t)r];
    std::vector<int> hb((size_t)nsl + 1, 0);
    for (int k = 1; k < nsl; ++k) {
        long long target = s->nnz_k * k / nsl;
        int r = 0;
        while (r < s->m_k && cnt[(size_t)r] < target) ++r;
        hb[(size_t)k] = std::max(r, hb[(size_t)k - 1]);
    }
    hb[(size_t)nsl] = s->m_k;
    return hb;
}

// The operator's layout pipeline (grid bricks or runs of TW columns, column
// order, ray sort, mode-2 tiles, mode-1 segment groups), with both self
// checks and a direct check of the ray permutation.
static void layouts(const plsqr_system *s, const plsqr_grid *grid, int nsl, int window, bool sort_damping,
                    bool too_many_segments = false) {
    std::vector<int> hb = slice_bounds(s, nsl);
    std::vector<int> perm, tiles;
    int tw = 128;
    if (grid) {
        tw = 256;
        plsqr::brick_order(grid->nx, grid->ny, grid->nz, 4, 8, 8, 1, perm, tiles);
    } else {
        perm.resize((size_t)s->n);
        for (int i = 0; i < s->n; ++i) perm[(size_t)i] = i;
        for (int c = 0; c < s->n; c += tw) tiles.push_back(c);
        tiles.push_back(s->n);
    }
    std::vector<int> seen((size_t)s->n, 0);
    for (int c = 0; c < s->n; ++c) seen[(size_t)perm[(size_t)c]]++;
    bool bij = true;
    for (int c = 0; c < s->n; ++c) bij = bij && seen[(size_t)c] == 1;
    CHECK(bij, "column permutation is not a bijection");
    plsqr::order_tile_columns(s, tiles, perm);
    plsqr::RaySorted rays;
    plsqr::sort_rays(s, perm, tiles, hb, grid, &rays, sort_damping);
    const plsqr_system *sys = &rays.sys;
    CHECK(sys->m_k == s->m_k && sys->nnz_k == s->nnz_k, "ray sort changed the kernel size");
    const size_t rcov = sort_damping ? (size_t)(s->m_k + s->m_d) : (size_t)s->m_k;
    CHECK(rays.rperm.size() >= rcov && rays.rinv.size() >= rcov, "row permutation too short");
    for (size_t r = 0; r < rcov && r < rays.rperm.size(); ++r)
        CHECK(rays.rinv[(size_t)rays.rperm[r]] == (int)r, "rinv[rperm[%zu]] != %zu", r, r);
    // the sorted kernel, mapped back, is the original operator
    std::vector<double> x((size_t)s->n);
    for (int c = 0; c < s->n; ++c) x[(size_t)c] = std::sin(0.37 * c + 1.0);
    std::vector<double> y0 = csc_apply(s, x), y1 = csc_apply(sys, x);
    double dmax = 0;
    for (int r = 0; r < s->m_k; ++r) dmax = std::fmax(dmax, std::fabs(y1[(size_t)rays.rinv[(size_t)r]] - y0[(size_t)r]));
    CHECK(dmax <= 1e-12, "ray-sorted kernel differs by %g", dmax);
    plsqr::TileLayout L;
    std::string why;
    bool ok = plsqr::build_tile_layout(sys, perm, tiles, hb, tw, &L, &why);
    if (too_many_segments) {
        // more than 65535 rays cross one brick of one slice: refused
        CHECK(!ok && !why.empty(), "an item with more than 65535 segments was accepted");
        std::printf("layout n=%d m_k=%d: refused (%s)\n", s->n, s->m_k, why.c_str());
        return;
    }
    CHECK(ok, "build_tile_layout: %s", why.c_str());
    if (!ok) return;
    long long cuts = L.cut_base.empty() ? 0 : L.cut_base.back();
    CHECK(plsqr::tile_layout_self_check(sys, L, &why), "tile_layout_self_check: %s", why.c_str());
    plsqr::RowTiles R;
    ok = plsqr::build_row_tiles(L, sys->m_k, &R, &why, window);
    CHECK(ok, "build_row_tiles: %s", why.c_str());
    if (!ok) return;
    CHECK(plsqr::row_tiles_self_check(sys, L, R, &why), "row_tiles_self_check: %s", why.c_str());
    if (window > 1) {
        // lanes sorted by length inside each window; seg_dst a permutation of
        // each item's ray-order indices; the same sums, bit for bit, as ray order
        const int G = plsqr::RowTiles::G;
        CHECK(R.seg_dst.size() == (size_t)R.n_grp * G, "seg_dst holds %zu lanes", R.seg_dst.size());
        for (long long it = 0; it < (long long)L.T * L.S; ++it) {
            const int ns = (int)(L.seg_start[(size_t)it + 1] - L.seg_start[(size_t)it]);
            std::vector<char> seen((size_t)ns, 0);
            for (int j = 0; j < ns; ++j) {
                const size_t lane = (size_t)R.grp_base[(size_t)it] * G + j;
                const int d = R.seg_dst[lane];
                CHECK(d < ns && !seen[(size_t)d], "item %lld: seg_dst %d repeated or out of range", it, d);
                if (d < ns) seen[(size_t)d] = 1;
                if (j % window != 0)
                    CHECK(R.seg_len[lane - 1] >= R.seg_len[lane], "item %lld: lengths not descending in a window", it);
            }
        }
        plsqr::RowTiles R1;
        CHECK(plsqr::build_row_tiles(L, sys->m_k, &R1, &why, 1) && R1.seg_dst.empty(), "ray-order layout");
        std::vector<double> xv((size_t)s->n), y1((size_t)sys->m_k), yw((size_t)sys->m_k);
        for (int c = 0; c < s->n; ++c) xv[(size_t)c] = std::sin(0.37 * c + 0.1);
        plsqr::row_tiles_apply(L, R1, sys->m_k, xv.data(), y1.data());
        plsqr::row_tiles_apply(L, R, sys->m_k, xv.data(), yw.data());
        CHECK(y1 == yw, "sorted windows change mode-1 sums");
        CHECK(R1.ray_seg == R.ray_seg, "sorted windows change the per-ray partial indices");
    }
    // compacted slot rows: the same sums, fewer slots, and the mode-1 builder refuses them
    const long long padded = L.n_slot;
    std::vector<double> u((size_t)s->m_k), P0((size_t)L.S * s->n), P1((size_t)L.S * s->n);
    for (int r = 0; r < s->m_k; ++r) u[(size_t)r] = std::cos(0.21 * r + 0.3);
    CHECK(plsqr::tile_layout_apply(L, s->n, u.data(), P0.data()), "padded apply");
    {
        // by position in TW-wide rows
        plsqr::TileLayout W = L;
        plsqr::compact_tile_layout(&W, false);
        CHECK(W.n_slot == padded && !W.packed, "position rows changed the slot count");
        CHECK(plsqr::tile_layout_apply(W, s->n, u.data(), P1.data()), "position-row apply: round cuts disagree");
        CHECK(P0 == P1, "slots by position give different sums");
    }
    plsqr::compact_tile_layout(&L);
    CHECK(plsqr::tile_layout_apply(L, s->n, u.data(), P1.data()), "compact apply: round cuts disagree");
    CHECK(P0 == P1, "compacted slots give different sums");
    CHECK(plsqr::tile_layout_self_check(sys, L, &why), "compact self check: %s", why.c_str());
    long long nnz_slots = 0;
    for (uint16_t l : L.vox_len) nnz_slots += l;
    CHECK(L.n_slot == nnz_slots && nnz_slots == sys->nnz_k, "compact form holds %lld slots for %lld nonzeros",
          L.n_slot, sys->nnz_k);
    plsqr::RowTiles R2;
    CHECK(!plsqr::build_row_tiles(L, sys->m_k, &R2, &why), "mode-1 layout built from compacted slots");
    std::printf("layout n=%d m_k=%d slices=%d window=%d grid=%d: %lld segments (at most %d per item), %lld slots, "
                "%lld round cuts, %d groups; compacted %lld -> %lld slots\n", s->n, s->m_k, L.S, window, grid ? 1 : 0,
                L.n_seg, L.max_seg_per_item, padded, cuts, R.n_grp, padded, L.n_slot);
}

// Mode 1 on super-bricks of 4 x 8 x 8 bricks: 1 x 1 x 2 of them (4 x 8 x
// 16, 512 voxels) number the voxels as brick order 1 does, 2 x 1 x 2 (8 x 8
// x 16, 1024 voxels) in an order of their own; either way every super-brick
// is a run of whole bricks, so the mode-1 tiles are unions of the 256-voxel
// tiles and one column order serves both modes.  Entries keep 8-bit voxels
// plus per-segment counts of the entries below voxel 256, 512, 768.
static void stacked(const plsqr_system *s, const plsqr_grid *grid, int nsl, int qx, int qz) {
    const int nx = grid->nx, ny = grid->ny, nz = grid->nz;
    std::vector<int> hb = slice_bounds(s, nsl);
    std::vector<int> perm, tiles, supers;
    plsqr::brick_order_grouped(nx, ny, nz, 4, 8, 8, qx, 1, qz, perm, tiles, supers);
    if (qx == 1) {
        std::vector<int> perm1, tiles1, perm2, tiles2;
        plsqr::brick_order(nx, ny, nz, 4, 8, 8, 1, perm1, tiles1);
        plsqr::brick_order(nx, ny, nz, 4, 8, 16, 1, perm2, tiles2);
        CHECK(perm == perm1 && perm == perm2 && tiles == tiles1 && supers == tiles2,
              "1 x 1 x 2 super-bricks number the columns differently from brick order 1");
    }
    // a permutation; every brick and super-brick a box of the stated extent
    std::vector<int> inv((size_t)s->n, -1);
    for (int c = 0; c < s->n; ++c) inv[(size_t)perm[(size_t)c]] = c;
    CHECK(std::find(inv.begin(), inv.end(), -1) == inv.end(), "grouped brick order is not a permutation");
    auto box = [&](const std::vector<int> &st, int ex, int ey, int ez, const char *what) {
        for (size_t t = 0; t + 1 < st.size(); ++t) {
            int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-1, -1, -1};
            for (int k = st[t]; k < st[t + 1]; ++k) {
                const int o = inv[(size_t)k], xyz[3] = {o % nx, (o / nx) % ny, o / (nx * ny)};
                for (int d = 0; d < 3; ++d) lo[d] = std::min(lo[d], xyz[d]), hi[d] = std::max(hi[d], xyz[d]);
            }
            const long long vol = (long long)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1);
            CHECK(hi[0] - lo[0] < ex && hi[1] - lo[1] < ey && hi[2] - lo[2] < ez && vol == st[t + 1] - st[t],
                  "%s %zu is not a box of at most %d x %d x %d voxels", what, t, ex, ey, ez);
        }
    };
    box(tiles, 4, 8, 8, "brick");
    box(supers, 4 * qx, 8, 8 * qz, "super-brick");
    bool subset = true;
    for (int b : supers) subset = subset && std::binary_search(tiles.begin(), tiles.end(), b);
    CHECK(subset, "a super-brick is not a union of brick tiles");
    plsqr::order_tile_columns(s, tiles, perm);
    plsqr::RaySorted rays;
    plsqr::sort_rays(s, perm, tiles, hb, grid, &rays);
    plsqr::TileLayout L;
    std::string why;
    const int TW = 256 * qx * qz;
    bool ok = plsqr::build_tile_layout(&rays.sys, perm, supers, hb, TW, &L, &why);
    CHECK(ok, "stacked build_tile_layout: %s", why.c_str());
    if (!ok) return;
    plsqr::RowTiles R;
    ok = plsqr::build_row_tiles(L, rays.sys.m_k, &R, &why);
    CHECK(ok, "stacked build_row_tiles: %s", why.c_str());
    if (!ok) return;
    CHECK(R.cut_bytes == (TW > 512 ? 4 : 1) && R.seg_lo.size() == (size_t)R.n_grp * plsqr::RowTiles::G * R.cut_bytes,
          "per-segment counts of %d bytes for %d-voxel tiles", R.cut_bytes, TW);
    // the counts: non-decreasing prefixes of each segment, none past its length
    for (int g = 0; g < R.n_grp; ++g)
        for (int l = 0; l < plsqr::RowTiles::G; ++l) {
            const uint8_t *c = &R.seg_lo[((size_t)g * plsqr::RowTiles::G + l) * R.cut_bytes];
            const int len = R.seg_len[(size_t)g * plsqr::RowTiles::G + l], nc = (TW - 1) / 256;
            if (len == 0) continue;  // no segment in this lane
            for (int j = 0; j < nc; ++j)
                CHECK(c[j] <= len && (j == 0 || c[j - 1] <= c[j]), "segment counts out of order");
            for (int j = nc; j < R.cut_bytes; ++j) CHECK(c[j] == 255, "unused count byte not 255");
        }
    CHECK(plsqr::row_tiles_self_check(&rays.sys, L, R, &why), "stacked row_tiles_self_check: %s", why.c_str());
    std::printf("mode-1 tiles of %d x 1 x %d bricks n=%d m_k=%d: %lld segments, %d groups\n", qx, qz, s->n, s->m_k,
                L.n_seg, R.n_grp);
}

static void files(const plsqr_system *s, const std::string &dir) {
    int rc = plsqr_write_reference_files(s, dir.c_str(), "k.bin", "k.info", "d.bin", "d.info", "dc.bin", "dc.info",
                                         "b.txt");
    CHECK(rc == PLSQR_OK, "write_reference_files %d", rc);
    plsqr_system t;
    std::memset(&t, 0, sizeof t);
    rc = plsqr_read_reference_files(dir.c_str(), "k.bin", "k.info", "d.bin", "d.info", "dc.bin", "dc.info", "b.txt",
                                    s->m_k, s->m_d, s->n, &t);
    CHECK(rc == PLSQR_OK, "read_reference_files %d: %s", rc, plsqr_last_global_error());
    if (rc != PLSQR_OK) return;
    CHECK(t.nnz_k == s->nnz_k && t.nnz_d == s->nnz_d, "nnz differ after the round trip");
    CHECK(!std::memcmp(t.k_col_ptr, s->k_col_ptr, sizeof(int) * ((size_t)s->n + 1)), "k_col_ptr differs");
    CHECK(!std::memcmp(t.k_row_idx, s->k_row_idx, sizeof(int) * (size_t)s->nnz_k), "k_row_idx differs");
    CHECK(!std::memcmp(t.k_val, s->k_val, sizeof(double) * (size_t)s->nnz_k), "k_val differs");
    CHECK(!std::memcmp(t.dc_row_idx, s->dc_row_idx, sizeof(int) * (size_t)s->nnz_d), "dc_row_idx differs");
    CHECK(!std::memcmp(t.d_val, s->d_val, sizeof(double) * (size_t)s->nnz_d), "d_val differs");
    CHECK(!std::memcmp(t.b, s->b, sizeof(double) * (size_t)s->m_k), "b differs");
    plsqr_system_free(&t);
    // a wrong row count is refused, not read past
    std::memset(&t, 0, sizeof t);
    rc = plsqr_read_reference_files(dir.c_str(), "k.bin", "k.info", "d.bin", "d.info", "dc.bin", "dc.info", "b.txt",
                                    s->m_k / 2, s->m_d, s->n, &t);
    CHECK(rc != PLSQR_OK, "a short row count was accepted");
    if (rc == PLSQR_OK) plsqr_system_free(&t);
    std::vector<double> x((size_t)s->n), x2((size_t)s->n + 1);
    for (int c = 0; c < s->n; ++c) x[(size_t)c] = 1.0 / (c + 1.0);
    CHECK(plsqr_write_x((dir + "/x.bin").c_str(), x.data(), s->n) == PLSQR_OK, "write_x");
    CHECK(plsqr_read_x((dir + "/x.bin").c_str(), x2.data(), s->n) == PLSQR_OK &&
              std::equal(x.begin(), x.end(), x2.begin()), "read_x");
    CHECK(plsqr_read_x((dir + "/x.bin").c_str(), x2.data(), s->n + 1) != PLSQR_OK, "read_x past the end accepted");
}

static void damping(const std::string &dir) {
    for (int cross = 0; cross < 2; ++cross) {
        plsqr_damping_spec sp;
        std::memset(&sp, 0, sizeof sp);
        sp.nx = 7;
        sp.ny = 5;
        sp.nz = 4;
        sp.n_param = 2;
        for (int i = 0; i < 7; ++i) sp.w[i] = 0.5 + 0.25 * i;
        sp.cross = cross;
        plsqr_system d;
        std::memset(&d, 0, sizeof d);
        int rc = plsqr_gen_damping(&sp, &d);
        CHECK(rc == PLSQR_OK, "gen_damping %d", rc);
        if (rc) continue;
        // the CSR and the CSC hold the same entries
        std::vector<double> dense((size_t)d.m_d * d.n, 0.0);
        for (int r = 0; r < d.m_d; ++r)
            for (int e = d.d_row_ptr[r]; e < d.d_row_ptr[r + 1]; ++e) dense[(size_t)r * d.n + d.d_col_idx[e]] += d.d_val[e];
        for (int c = 0; c < d.n; ++c)
            for (int e = d.dc_col_ptr[c]; e < d.dc_col_ptr[c + 1]; ++e) dense[(size_t)d.dc_row_idx[e] * d.n + c] -= d.dc_val[e];
        double dm = 0;
        for (double v : dense) dm = std::fmax(dm, std::fabs(v));
        CHECK(dm == 0.0, "damping CSR and CSC differ by %g", dm);
        rc = plsqr_write_damping_files(&sp, dir.c_str(), "dd.bin", "dd.info", "ddc.bin", "ddc.info");
        CHECK(rc == PLSQR_OK, "write_damping_files %d", rc);
        std::printf("damping cross=%d: %d rows, %lld entries\n", cross, d.m_d, d.nnz_d);
        plsqr_system_free(&d);
    }
}

static void ascii_rows(const std::string &dir) {
    const std::string in = dir + "/ascii";
    std::string cmd = "mkdir -p '" + in + "'";
    CHECK(std::system(cmd.c_str()) == 0, "mkdir");
    FILE *l = std::fopen((dir + "/ascii.list").c_str(), "w");
    std::fprintf(l, "3\n");
    for (int r = 0; r < 3; ++r) {
        std::fprintf(l, "row%d\n", r);
        FILE *f = std::fopen((in + "/row" + std::to_string(r)).c_str(), "w");
        for (int k = 0; k <= 4 * r; ++k) std::fprintf(f, "%d %d %d %d %.17e\n", 3 * k + r + 1, k, r, 1, 1e-6 * (k - 2.5));
        std::fclose(f);
    }
    std::fclose(l);
    int rc = plsqr_ker_ascii2bin((dir + "/ascii.list").c_str(), in.c_str(), dir.c_str(), "bin.list");
    CHECK(rc == PLSQR_OK, "ker_ascii2bin %d: %s", rc, plsqr_last_global_error());
    FILE *f = std::fopen((dir + "/row2.bin").c_str(), "rb");
    CHECK(f, "row2.bin missing");
    if (f) {
        std::fseek(f, 0, SEEK_END);
        CHECK(std::ftell(f) == 9 * 16, "row2.bin holds %ld bytes", std::ftell(f));
        std::fclose(f);
    }
}

static void batch() {
    const int P = 3;
    plsqr_system sy[P];
    plsqr_grid gr[P] = {{9, 8, 6}, {12, 10, 7}, {0, 0, 0}};
    plsqr_gen_params ps[P] = {params(9, 8, 6, 400, 1, 11), params(12, 10, 7, 900, 0, 12),
                              params(8, 8, 8, 300, 1, 13)};
    const plsqr_system *ptr[P];
    for (int p = 0; p < P; ++p) {
        std::memset(&sy[p], 0, sizeof sy[p]);
        CHECK(plsqr_gen_system(&ps[p], &sy[p]) == PLSQR_OK, "gen_system %d", p);
        std::string err;
        CHECK(plsqr::validate_system(&sy[p], &err) == PLSQR_OK, "validate_system: %s", err.c_str());
        ptr[p] = &sy[p];
    }
    plsqr::BatchPlan plan;
    std::string err;
    int rc = plsqr::build_batch_plan(ptr, gr, P, 128, 512, 2048, 1024, 256, 8, &plan, &err);
    CHECK(rc == PLSQR_OK, "build_batch_plan %d: %s", rc, err.c_str());
    if (rc == PLSQR_OK) {
        long long rows = 0, cols = 0;
        for (int p = 0; p < P; ++p) {
            rows += sy[p].m_k + sy[p].m_d;
            cols += sy[p].n;
        }
        CHECK(plan.rows == rows && plan.cols == cols, "plan covers %lld rows / %lld cols", plan.rows, plan.cols);
        for (int b : plan.blk_item) CHECK(b >= -1 && b < plan.items, "blk_item %d out of range", b);
        std::printf("batch: %d items, %lld segments, %lld groups\n", plan.items, plan.segs, plan.groups);
    }
    // a broken system is reported, not read past
    sy[0].k_row_idx[0] = sy[0].m_k + 5;
    CHECK(plsqr::validate_system(&sy[0], &err) == PLSQR_ERR_BOUNDS, "out-of-range row accepted");
    sy[0].k_row_idx[0] = 0;
    for (int p = 0; p < P; ++p) plsqr_system_free(&sy[p]);
}

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: host_layouts <scratch dir>\n");
        return 2;
    }
    const std::string dir = argv[1];
    // brick grid (top sources), grid with a ragged last brick (opposite
    // faces), no grid, more columns than rows, items of several rounds
    // (more than 1024 segments), and one brick crossed by too many rays
    struct Case { int nx, ny, nz; long long rays; int geom; int slices; bool grid; bool refused; };
    const Case cases[] = {{16, 16, 16, 6000, 1, 8, true, false}, {13, 11, 9, 3000, 0, 4, true, false},
                          {10, 10, 10, 2500, 1, 1, false, false}, {20, 18, 16, 300, 1, 1, true, false},
                          {8, 8, 8, 6000, 1, 1, true, false}, {4, 8, 8, 70000, 1, 1, true, true}};
    for (const Case &c : cases) {
        plsqr_gen_params p = params(c.nx, c.ny, c.nz, c.rays, c.geom, 1234 + c.nx);
        plsqr_system s;
        std::memset(&s, 0, sizeof s);
        int rc = plsqr_gen_system(&p, &s);
        CHECK(rc == PLSQR_OK, "gen_system %d", rc);
        if (rc) continue;
        plsqr_grid g = {c.nx, c.ny, c.nz};
        if (c.refused) {
            layouts(&s, &g, c.slices, 1, true, true);
            plsqr_system_free(&s);
            continue;
        }
        for (int window : {1, 64, 1024}) layouts(&s, c.grid ? &g : nullptr, c.slices, window, true);
        layouts(&s, c.grid ? &g : nullptr, c.slices, 1, false);
        if (c.grid) {
            stacked(&s, &g, c.slices, 1, 2);
            stacked(&s, &g, c.slices, 2, 2);
            stacked(&s, &g, c.slices, 2, 1);
        }
        if (&c == &cases[0]) files(&s, dir);
        plsqr_system_free(&s);
    }
    damping(dir);
    ascii_rows(dir);
    batch();
    if (failures) {
        std::printf("%d failures\n", failures);
        return 1;
    }
    std::printf("host layouts ok\n");
    return 0;
}
