// Host-side layout of a batch of independent systems; see batch_build.h.
#include "batch_build.h"

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstdlib>

namespace plsqr {

namespace {

int check_ptr(const char *what, const int *p, int len, long long nnz, std::string *err) {
    if (!p) {
        *err = std::string(what) + " is NULL";
        return PLSQR_ERR_ARG;
    }
    if (p[0] != 0) {
        *err = std::string(what) + "[0] != 0";
        return PLSQR_ERR_FORMAT;
    }
    for (int i = 0; i < len; ++i)
        if (p[i + 1] < p[i]) {
            *err = std::string(what) + " decreases at " + std::to_string(i);
            return PLSQR_ERR_FORMAT;
        }
    if (p[len] != nnz) {
        *err = std::string(what) + " does not end at nnz";
        return PLSQR_ERR_FORMAT;
    }
    return PLSQR_OK;
}

int check_idx(const char *what, const int *idx, const double *val, long long nnz, int bound, std::string *err) {
    if (nnz > 0 && (!idx || !val)) {
        *err = std::string(what) + " arrays are NULL";
        return PLSQR_ERR_ARG;
    }
    std::atomic<long long> bad_idx{LLONG_MAX}, bad_val{LLONG_MAX};
    par_for(nnz, [&](long long b, long long e) {
        for (long long i = b; i < e; ++i) {
            if (idx[i] < 0 || idx[i] >= bound) {
                long long cur = bad_idx.load();
                while (i < cur && !bad_idx.compare_exchange_weak(cur, i)) {
                }
                return;
            }
            if (!std::isfinite(val[i])) {
                long long cur = bad_val.load();
                while (i < cur && !bad_val.compare_exchange_weak(cur, i)) {
                }
                return;
            }
        }
    });
    if (bad_idx.load() != LLONG_MAX) {
        *err = std::string(what) + " index out of range at entry " + std::to_string(bad_idx.load());
        return PLSQR_ERR_BOUNDS;
    }
    if (bad_val.load() != LLONG_MAX) {
        *err = std::string(what) + " value not finite at entry " + std::to_string(bad_val.load());
        return PLSQR_ERR_NONFINITE;
    }
    return PLSQR_OK;
}

// CSR of [A_k ; D] in the new column numbering.  Kernel rows list their
// columns ascending (the CSC is walked in new column order); damping rows keep
// their entry order.  Row ranges are filled by independent host threads.
void build_csr(const plsqr_system *s, const std::vector<int> &perm, const std::vector<int> &inv, ProblemLayout *L) {
    const int m_k = s->m_k, m_d = s->m_d, n = s->n;
    L->rp.assign((size_t)m_k + m_d + 1, 0);
    std::vector<int> cnt((size_t)m_k, 0);
    for (long long e = 0; e < s->nnz_k; ++e) ++cnt[(size_t)s->k_row_idx[e]];
    for (int r = 0; r < m_k; ++r) L->rp[(size_t)r + 1] = L->rp[(size_t)r] + cnt[(size_t)r];
    for (int r = 0; r < m_d; ++r)
        L->rp[(size_t)m_k + r + 1] = L->rp[(size_t)m_k + r] + (s->d_row_ptr[r + 1] - s->d_row_ptr[r]);
    const long long nnz = (long long)L->rp[(size_t)m_k + m_d];
    L->ci.assign((size_t)nnz, 0);
    L->rv.assign((size_t)nnz, 0.0);
    par_for(m_k, [&](long long r0, long long r1) {
        std::vector<int> next(L->rp.begin() + r0, L->rp.begin() + r1);
        for (int c = 0; c < n; ++c) {
            const int o = inv[(size_t)c];
            const int *rb = s->k_row_idx + s->k_col_ptr[o];
            const int *re = s->k_row_idx + s->k_col_ptr[o + 1];
            const int *lo = std::lower_bound(rb, re, (int)r0);
            for (const int *q = lo; q != re && *q < r1; ++q) {
                const int pos = next[(size_t)(*q - r0)]++;
                L->ci[(size_t)pos] = c;
                L->rv[(size_t)pos] = s->k_val[q - s->k_row_idx];
            }
        }
    });
    for (int r = 0; r < m_d; ++r) {
        int pos = L->rp[(size_t)m_k + r];
        for (int e = s->d_row_ptr[r]; e < s->d_row_ptr[r + 1]; ++e, ++pos) {
            L->ci[(size_t)pos] = perm[(size_t)s->d_col_idx[e]];
            L->rv[(size_t)pos] = s->d_val[e];
        }
    }
    // damping columns in the new order
    L->dcp.assign((size_t)n + 1, 0);
    for (int c = 0; c < n; ++c) {
        const int o = inv[(size_t)c];
        L->dcp[(size_t)c + 1] = L->dcp[(size_t)c] + (s->dc_col_ptr[o + 1] - s->dc_col_ptr[o]);
    }
    L->dri.assign((size_t)L->dcp[(size_t)n], 0);
    L->dcv.assign((size_t)L->dcp[(size_t)n], 0.0);
    for (int c = 0; c < n; ++c) {
        const int o = inv[(size_t)c];
        int pos = L->dcp[(size_t)c];
        for (int e = s->dc_col_ptr[o]; e < s->dc_col_ptr[o + 1]; ++e, ++pos) {
            L->dri[(size_t)pos] = s->dc_row_idx[e];
            L->dcv[(size_t)pos] = s->dc_val[e];
        }
    }
}

}  // namespace

int validate_system(const plsqr_system *s, std::string *err) {
    if (!s) {
        *err = "system is NULL";
        return PLSQR_ERR_ARG;
    }
    if (s->m_k <= 0 || s->n <= 0 || s->m_d < 0 || s->nnz_k < 0 || s->nnz_d < 0) {
        *err = "system dimensions must be positive";
        return PLSQR_ERR_ARG;
    }
    if (s->nnz_k > INT_MAX || s->nnz_d > INT_MAX) {
        *err = "more than 2^31-1 nonzeros in one system";
        return PLSQR_ERR_ARG;
    }
    int rc;
    if ((rc = check_ptr("kernel col_ptr", s->k_col_ptr, s->n, s->nnz_k, err)) ||
        (rc = check_idx("kernel row_idx", s->k_row_idx, s->k_val, s->nnz_k, s->m_k, err)))
        return rc;
    for (int c = 0; c < s->n; ++c)
        for (int e = s->k_col_ptr[c] + 1; e < s->k_col_ptr[c + 1]; ++e)
            if (s->k_row_idx[e] <= s->k_row_idx[e - 1]) {
                *err = "kernel rows not ascending in column " + std::to_string(c);
                return PLSQR_ERR_FORMAT;
            }
    if (s->m_d > 0) {
        if ((rc = check_ptr("damping row_ptr", s->d_row_ptr, s->m_d, s->nnz_d, err)) ||
            (rc = check_idx("damping col_idx", s->d_col_idx, s->d_val, s->nnz_d, s->n, err)) ||
            (rc = check_ptr("damping col_ptr", s->dc_col_ptr, s->n, s->nnz_d, err)) ||
            (rc = check_idx("damping-by-column row_idx", s->dc_row_idx, s->dc_val, s->nnz_d, s->m_d, err)))
            return rc;
    } else if (s->nnz_d != 0) {
        *err = "damping nonzeros without damping rows";
        return PLSQR_ERR_FORMAT;
    }
    return PLSQR_OK;
}

int build_batch_plan(const plsqr_system *const *systems, const plsqr_grid *grids, int P, int TW, int row_chunk,
                     int damp_chunk, int col_chunk, int slice_rows, int xcds, BatchPlan *plan, std::string *err) {
    if (P <= 0 || !systems) {
        *err = "no systems";
        return PLSQR_ERR_ARG;
    }
    plan->P = P;
    plan->TW = TW;
    plan->prob.assign((size_t)P, ProblemLayout());
    int rc;
    for (int p = 0; p < P; ++p) {
        const plsqr_system *s = systems[p];
        if ((rc = validate_system(s, err))) {
            *err = "system " + std::to_string(p) + ": " + *err;
            return rc;
        }
        const plsqr_grid *g = grids ? &grids[p] : nullptr;
        if (g && g->nx * (long long)g->ny * g->nz == 0) g = nullptr;
        if (g && (long long)g->nx * g->ny * g->nz != s->n) {
            *err = "system " + std::to_string(p) + ": grid dimensions do not multiply to the number of columns";
            return PLSQR_ERR_ARG;
        }
        ProblemLayout &L = plan->prob[(size_t)p];
        L.m_k = s->m_k;
        L.m_d = s->m_d;
        L.n = s->n;
        std::vector<int> perm, tiles;
        if (g) {
            brick_order(g->nx, g->ny, g->nz, 4, 4, 8, std::getenv("PLSQR_BRICKS") ? std::atoi(std::getenv("PLSQR_BRICKS")) : 1, perm, tiles);
        } else {
            perm.resize((size_t)s->n);
            for (int i = 0; i < s->n; ++i) perm[(size_t)i] = i;
            for (int c = 0; c < s->n; c += TW) tiles.push_back(c);
            tiles.push_back(s->n);
        }
        order_tile_columns(s, tiles, perm);
        // row slices balanced by nonzeros, about slice_rows rows each
        const int S = std::max(1, std::min(64, (s->m_k + std::max(slice_rows, 1) - 1) / std::max(slice_rows, 1)));
        auto cut_slices = [S](const plsqr_system *q_s) {
            std::vector<int> sl((size_t)S + 1, 0);
            std::vector<long long> cnt((size_t)q_s->m_k + 1, 0);
            for (long long e = 0; e < q_s->nnz_k; ++e) cnt[(size_t)q_s->k_row_idx[e] + 1]++;
            for (int r = 0; r < q_s->m_k; ++r) cnt[(size_t)r + 1] += cnt[(size_t)r];
            for (int q = 1; q < S; ++q) {
                const long long target = q_s->nnz_k * q / S;
                int b = (int)(std::lower_bound(cnt.begin(), cnt.end(), target) - cnt.begin());
                sl[(size_t)q] = std::min(std::max(b, sl[(size_t)q - 1]), q_s->m_k);
            }
            sl[(size_t)S] = q_s->m_k;
            return sl;
        };
        // rays sorted by (first tile, last tile): within the slices (default), or
        // over all rays with the slices cut afterwards on the sorted rows
        const bool global_sort = std::getenv("PLSQR_BATCH_RAYSORT") && std::atoi(std::getenv("PLSQR_BATCH_RAYSORT")) == 2;
        std::vector<int> slices = global_sort ? std::vector<int>() : cut_slices(s);
        L.S = S;
        RaySorted rays;
        // damping rows stay in place: the batch keeps its own damping row chunks
        sort_rays(s, perm, tiles, slices, (std::getenv("PLSQR_RAYKEY") && std::atoi(std::getenv("PLSQR_RAYKEY")) == 0) ? nullptr : g, &rays,
                  false);
        L.rperm = rays.rperm;
        const plsqr_system *ss = &rays.sys;
        if (global_sort) slices = cut_slices(ss);
        if (!build_tile_layout(ss, perm, tiles, slices, TW, &L.tiles, err)) {
            *err = "system " + std::to_string(p) + ": " + *err;
            return PLSQR_ERR_FORMAT;
        }
        if (std::getenv("PLSQR_TILE_CHECK") && std::atoi(std::getenv("PLSQR_TILE_CHECK")) != 0 &&
            !tile_layout_self_check(ss, L.tiles, err)) {
            *err = "system " + std::to_string(p) + ": " + *err;
            return PLSQR_ERR_STATE;
        }
        L.inv = L.tiles.inv;
        L.tiles.perm.clear();
        L.tiles.perm.shrink_to_fit();
        build_csr(ss, perm, L.inv, &L);
    }
    // offsets and pieces
    for (auto *v : {&plan->row0, &plan->col0, &plan->nnz0, &plan->dnnz0, &plan->slot0, &plan->seg0, &plan->cut0,
                    &plan->pofs, &plan->grp0, &plan->ent0})
        v->assign((size_t)P + 1, 0);
    plan->item0.assign((size_t)P + 1, 0);
    plan->nslice.assign((size_t)P, 1);
    plan->rch0.assign((size_t)P + 1, 0);
    plan->cch0.assign((size_t)P + 1, 0);
    plan->rch_row.clear();
    plan->rch_prob.clear();
    plan->rch_kind.clear();
    plan->cch_col.clear();
    plan->cch_prob.clear();
    int mx = 0;
    std::vector<int> rch_slice;  // slice of every row chunk (-1: damping rows)
    for (int p = 0; p < P; ++p) {
        const ProblemLayout &L = plan->prob[(size_t)p];
        const long long r0 = plan->row0[(size_t)p], c0 = plan->col0[(size_t)p];
        plan->row0[(size_t)p + 1] = r0 + L.m_k + L.m_d;
        plan->col0[(size_t)p + 1] = c0 + L.n;
        plan->nnz0[(size_t)p + 1] = plan->nnz0[(size_t)p] + (long long)L.rp.back();
        plan->dnnz0[(size_t)p + 1] = plan->dnnz0[(size_t)p] + (long long)L.dcp.back();
        plan->slot0[(size_t)p + 1] = plan->slot0[(size_t)p] + L.tiles.n_slot;
        plan->seg0[(size_t)p + 1] = plan->seg0[(size_t)p] + L.tiles.n_seg;
        plan->cut0[(size_t)p + 1] = plan->cut0[(size_t)p] + L.tiles.cut_base.back();
        {
            // the mode-1 layout built from these tiles at upload: one group per
            // 64 segments of an item, one entry per kernel nonzero
            long long g = 0, e = 0;
            const long long nit = (long long)L.tiles.T * L.tiles.S;
            for (long long it = 0; it < nit; ++it)
                g += (L.tiles.seg_start[(size_t)it + 1] - L.tiles.seg_start[(size_t)it] + RowTiles::G - 1) / RowTiles::G;
            for (uint16_t l : L.tiles.vox_len) e += l;
            plan->grp0[(size_t)p + 1] = plan->grp0[(size_t)p] + g;
            plan->ent0[(size_t)p + 1] = plan->ent0[(size_t)p] + e;
        }
        plan->item0[(size_t)p + 1] = plan->item0[(size_t)p] + L.tiles.T * L.S;
        plan->pofs[(size_t)p + 1] = plan->pofs[(size_t)p] + (long long)L.S * L.n;
        plan->nslice[(size_t)p] = L.S;
        mx = std::max(mx, L.tiles.max_seg_per_item);
        // kernel-row chunks inside the row slices, so that every chunk's rays
        // belong to one slice (whose XCD holds their segment partials)
        {
            std::vector<int> sb = L.tiles.slice_bounds;
            if (sb.size() < 2) sb = {0, L.m_k};
            for (size_t q = 0; q + 1 < sb.size(); ++q)
                for (int r = sb[q]; r < sb[q + 1]; r += row_chunk) {
                    plan->rch_row.push_back((int)(r0 + r));
                    plan->rch_prob.push_back(p);
                    plan->rch_kind.push_back(0);
                    rch_slice.push_back((int)q);
                }
        }
        for (int r = 0; r < L.m_d; r += damp_chunk) {
            plan->rch_row.push_back((int)(r0 + L.m_k + r));
            plan->rch_prob.push_back(p);
            plan->rch_kind.push_back(1);
            rch_slice.push_back(-1);
        }
        plan->rch0[(size_t)p + 1] = (int)plan->rch_prob.size();
        for (int c = 0; c < L.n; c += col_chunk) {
            plan->cch_col.push_back((int)(c0 + c));
            plan->cch_prob.push_back(p);
        }
        plan->cch0[(size_t)p + 1] = (int)plan->cch_prob.size();
    }
    // slices to XCD queues, largest first to the least loaded queue; inside a
    // queue the items of one slice are consecutive (tile order)
    {
        struct Sl {
            int p, s;
            long long work;
        };
        std::vector<Sl> sl;
        for (int p = 0; p < P; ++p) {
            const ProblemLayout &L = plan->prob[(size_t)p];
            for (int q = 0; q < L.S; ++q) {
                long long w = 0;
                for (int t = 0; t < L.tiles.T; ++t) {
                    const long long it = (long long)t * L.S + q;
                    w += L.tiles.slot_base[(size_t)it + 1] - L.tiles.slot_base[(size_t)it] + 4 * TW;
                }
                sl.push_back({p, q, w});
            }
        }
        std::stable_sort(sl.begin(), sl.end(), [](const Sl &a, const Sl &b) { return a.work > b.work; });
        const int X = std::max(1, xcds);
        std::vector<long long> load((size_t)X, 0);
        std::vector<std::vector<int>> queue((size_t)X);
        for (const Sl &q : sl) {
            const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            load[(size_t)x] += q.work;
            const ProblemLayout &L = plan->prob[(size_t)q.p];
            for (int t = 0; t < L.tiles.T; ++t)
                queue[(size_t)x].push_back((int)(plan->item0[(size_t)q.p] + (long long)t * L.S + q.s));
        }
        size_t len = 0;
        for (const auto &v : queue) len = std::max(len, v.size());
        plan->blk_item.assign(len * X, -1);
        for (int x = 0; x < X; ++x)
            for (size_t i = 0; i < queue[(size_t)x].size(); ++i) plan->blk_item[i * X + x] = queue[(size_t)x][i];
        // row chunks: a slice's kernel-row chunks to its XCD's queue, damping
        // chunks to the queue with the fewest rows
        std::vector<std::vector<int>> xcd_of((size_t)P);
        for (int p = 0; p < P; ++p) xcd_of[(size_t)p].assign((size_t)plan->prob[(size_t)p].S, 0);
        {
            std::vector<long long> ld((size_t)X, 0);
            for (const Sl &q : sl) {
                const int x = (int)(std::min_element(ld.begin(), ld.end()) - ld.begin());
                ld[(size_t)x] += q.work;
                xcd_of[(size_t)q.p][(size_t)q.s] = x;
            }
        }
        std::vector<std::vector<int>> rq((size_t)X);
        std::vector<long long> rows_in((size_t)X, 0);
        const int nch = (int)plan->rch_prob.size();
        for (int ch = 0; ch < nch; ++ch) {
            const long long nr = (ch + 1 < nch ? plan->rch_row[(size_t)ch + 1]
                                               : plan->row0[(size_t)plan->rch_prob[(size_t)ch] + 1]) -
                                 plan->rch_row[(size_t)ch];
            int x;
            if (rch_slice[(size_t)ch] >= 0)
                x = xcd_of[(size_t)plan->rch_prob[(size_t)ch]][(size_t)rch_slice[(size_t)ch]];
            else
                x = (int)(std::min_element(rows_in.begin(), rows_in.end()) - rows_in.begin());
            rq[(size_t)x].push_back(ch);
            rows_in[(size_t)x] += nr;
        }
        size_t rlen = 0;
        for (const auto &v : rq) rlen = std::max(rlen, v.size());
        plan->blk_rch.assign(rlen * X, -1);
        for (int x = 0; x < X; ++x)
            for (size_t i = 0; i < rq[(size_t)x].size(); ++i) plan->blk_rch[i * X + x] = rq[(size_t)x][i];
    }
    plan->rows = plan->row0[(size_t)P];
    plan->cols = plan->col0[(size_t)P];
    plan->nnz = plan->nnz0[(size_t)P];
    plan->dnnz = plan->dnnz0[(size_t)P];
    plan->slots = plan->slot0[(size_t)P];
    plan->segs = plan->seg0[(size_t)P];
    plan->groups = plan->grp0[(size_t)P];
    plan->ents = plan->ent0[(size_t)P];
    plan->cuts = plan->cut0[(size_t)P];
    plan->items = plan->item0[(size_t)P];
    plan->partials = plan->pofs[(size_t)P];
    plan->max_seg_per_item = mx;
    plan->rch_row.push_back((int)plan->rows);
    plan->cch_col.push_back((int)plan->cols);
    if (plan->rows * 4 > INT_MAX || plan->cols * 4 > INT_MAX || plan->nnz > INT_MAX || plan->dnnz > INT_MAX ||
        plan->items >= INT_MAX / TW) {
        *err = "the batch exceeds 32-bit row, column or nonzero indexing";
        return PLSQR_ERR_ARG;
    }
    return PLSQR_OK;
}

}  // namespace plsqr
