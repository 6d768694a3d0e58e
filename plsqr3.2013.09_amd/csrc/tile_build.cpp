// Host-side construction of the tiled mode-2 layout; see tile_build.h.
#include "tile_build.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <atomic>
#include <cstdlib>
#include <thread>
#include <tuple>

namespace plsqr {

namespace {

// rows of old column o inside [r0, r1): index range in the CSC arrays
inline void slice_range(const plsqr_system *s, int o, int r0, int r1, int *b, int *e) {
    const int *rb = s->k_row_idx + s->k_col_ptr[o];
    const int *re = s->k_row_idx + s->k_col_ptr[o + 1];
    *b = (int)(std::lower_bound(rb, re, r0) - s->k_row_idx);
    *e = (int)(std::lower_bound(rb, re, r1) - s->k_row_idx);
}

}  // namespace

void brick_order(int nx, int ny, int nz, int bx, int by, int bz, int order,
                 std::vector<int> &perm, std::vector<int> &tile_start) {
    const long long n = (long long)nx * ny * nz;
    perm.assign((size_t)n, 0);
    tile_start.clear();
    int next = 0;
    auto brick = [&](int x0, int y0, int z0) {
        tile_start.push_back(next);
        for (int z = z0; z < std::min(nz, z0 + bz); ++z)
            for (int y = y0; y < std::min(ny, y0 + by); ++y)
                for (int x = x0; x < std::min(nx, x0 + bx); ++x)
                    perm[(size_t)z * ny * nx + (size_t)y * nx + x] = next++;
    };
    if (order == 2) {
        // Morton (Z-order) of the brick coordinates
        const int nbx = (nx + bx - 1) / bx, nby = (ny + by - 1) / by, nbz = (nz + bz - 1) / bz;
        std::vector<std::pair<uint64_t, int>> key;
        for (int z = 0; z < nbz; ++z)
            for (int y = 0; y < nby; ++y)
                for (int x = 0; x < nbx; ++x) {
                    uint64_t m = 0;
                    for (int i = 0; i < 20; ++i)
                        m |= (uint64_t)((x >> i) & 1) << (3 * i) | (uint64_t)((y >> i) & 1) << (3 * i + 1) |
                             (uint64_t)((z >> i) & 1) << (3 * i + 2);
                    key.emplace_back(m, (z * nby + y) * nbx + x);
                }
        std::sort(key.begin(), key.end());
        for (const auto &k : key) {
            const int id = k.second, x = id % nbx, y = (id / nbx) % nby, z = id / (nbx * nby);
            brick(x * bx, y * by, z * bz);
        }
    } else if (order == 1) {
        for (int y0 = 0; y0 < ny; y0 += by)
            for (int x0 = 0; x0 < nx; x0 += bx)
                for (int z0 = 0; z0 < nz; z0 += bz) brick(x0, y0, z0);
    } else {
        for (int z0 = 0; z0 < nz; z0 += bz)
            for (int y0 = 0; y0 < ny; y0 += by)
                for (int x0 = 0; x0 < nx; x0 += bx) brick(x0, y0, z0);
    }
    tile_start.push_back(next);
}

void brick_order_grouped(int nx, int ny, int nz, int bx, int by, int bz, int qx, int qy, int qz,
                         std::vector<int> &perm, std::vector<int> &tile_start, std::vector<int> &super_start) {
    const long long n = (long long)nx * ny * nz;
    perm.assign((size_t)n, 0);
    tile_start.clear();
    super_start.clear();
    int next = 0;
    const int sx = bx * qx, sy = by * qy, sz = bz * qz;  // super-brick extent in voxels
    for (int y1 = 0; y1 < ny; y1 += sy)
        for (int x1 = 0; x1 < nx; x1 += sx)
            for (int z1 = 0; z1 < nz; z1 += sz) {
                super_start.push_back(next);
                for (int y0 = y1; y0 < std::min(ny, y1 + sy); y0 += by)
                    for (int x0 = x1; x0 < std::min(nx, x1 + sx); x0 += bx)
                        for (int z0 = z1; z0 < std::min(nz, z1 + sz); z0 += bz) {
                            tile_start.push_back(next);
                            for (int z = z0; z < std::min(nz, z0 + bz); ++z)
                                for (int y = y0; y < std::min(ny, y0 + by); ++y)
                                    for (int x = x0; x < std::min(nx, x0 + bx); ++x)
                                        perm[(size_t)z * ny * nx + (size_t)y * nx + x] = next++;
                        }
            }
    tile_start.push_back(next);
    super_start.push_back(next);
}

bool build_tile_layout(const plsqr_system *s, const std::vector<int> &perm,
                       const std::vector<int> &tile_start, const std::vector<int> &slice_bounds,
                       int TW, TileLayout *L, std::string *err) {
    const int n = s->n;
    L->TW = TW;
    L->perm = perm;
    L->inv.assign((size_t)n, 0);
    for (int o = 0; o < n; ++o) L->inv[(size_t)perm[(size_t)o]] = o;
    L->tile_start = tile_start;
    L->slice_bounds = slice_bounds;
    L->T = (int)tile_start.size() - 1;
    L->S = (int)slice_bounds.size() - 1;
    const int T = L->T, S = L->S;
    const long long items = (long long)T * S;
    for (int t = 0; t < T; ++t)
        if (tile_start[(size_t)t + 1] - tile_start[(size_t)t] > TW || tile_start[(size_t)t + 1] < tile_start[(size_t)t]) {
            *err = "tile wider than the kernel's tile width";
            return false;
        }
    L->vox_len.assign((size_t)items * TW, 0);
    std::vector<int> nseg((size_t)items, 0), depth((size_t)items, 0);
    std::vector<std::vector<int>> rows_of((size_t)items);
    std::atomic<bool> ok{true};
    // pass 1: segments (distinct rows) and voxel lengths per item
    par_for(T, [&](long long tb, long long te) {
        std::vector<int> buf;
        for (long long t = tb; t < te; ++t) {
            const int c0 = tile_start[(size_t)t], tw = tile_start[(size_t)t + 1] - c0;
            for (int sl = 0; sl < S; ++sl) {
                const long long it = t * S + sl;
                buf.clear();
                int dmax = 0;
                for (int v = 0; v < tw; ++v) {
                    int b, e;
                    slice_range(s, L->inv[(size_t)(c0 + v)], slice_bounds[(size_t)sl], slice_bounds[(size_t)sl + 1], &b, &e);
                    const int len = e - b;
                    if (len > 65535) ok.store(false);
                    L->vox_len[(size_t)it * TW + v] = (uint16_t)std::min(len, 65535);
                    dmax = std::max(dmax, len);
                    buf.insert(buf.end(), s->k_row_idx + b, s->k_row_idx + e);
                }
                std::sort(buf.begin(), buf.end());
                buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
                nseg[(size_t)it] = (int)buf.size();
                depth[(size_t)it] = dmax;
                rows_of[(size_t)it] = buf;
            }
        }
    });
    if (!ok.load()) {
        *err = "a voxel holds more than 65535 nonzeros in one slice";
        return false;
    }
    L->seg_start.assign((size_t)items + 1, 0);
    L->slot_base.assign((size_t)items + 1, 0);
    L->cut_base.assign((size_t)items + 1, 0);
    int mx = 0;
    for (long long it = 0; it < items; ++it) {
        L->seg_start[(size_t)it + 1] = L->seg_start[(size_t)it] + nseg[(size_t)it];
        L->slot_base[(size_t)it + 1] = L->slot_base[(size_t)it] + (long long)depth[(size_t)it] * TW;
        const int ncut = nseg[(size_t)it] > 0 ? (nseg[(size_t)it] - 1) / TileLayout::CUT : 0;
        L->cut_base[(size_t)it + 1] = L->cut_base[(size_t)it] + ncut;
        mx = std::max(mx, nseg[(size_t)it]);
    }
    L->vox_cut.assign((size_t)L->cut_base[(size_t)items] * TW, 0);
    L->max_seg_per_item = mx;
    if (mx > 65535) {
        *err = "an item needs more than 65535 segments";
        return false;
    }
    L->n_seg = L->seg_start[(size_t)items];
    L->n_slot = L->slot_base[(size_t)items];
    L->seg_ray.assign((size_t)L->n_seg, 0);
    L->sval.assign((size_t)L->n_slot, 0.0);
    L->sseg.assign((size_t)L->n_slot, 0);
    // pass 2: fill segments and slots
    par_for(T, [&](long long tb, long long te) {
        for (long long t = tb; t < te; ++t) {
            const int c0 = tile_start[(size_t)t], tw = tile_start[(size_t)t + 1] - c0;
            for (int sl = 0; sl < S; ++sl) {
                const long long it = t * S + sl;
                const std::vector<int> &rows = rows_of[(size_t)it];
                std::copy(rows.begin(), rows.end(), L->seg_ray.begin() + L->seg_start[(size_t)it]);
                const long long base = L->slot_base[(size_t)it];
                for (int v = 0; v < tw; ++v) {
                    int b, e;
                    slice_range(s, L->inv[(size_t)(c0 + v)], slice_bounds[(size_t)sl], slice_bounds[(size_t)sl + 1], &b, &e);
                    const long long cb = L->cut_base[(size_t)it], ncut = L->cut_base[(size_t)it + 1] - cb;
                    long long q = 0;
                    for (int k = b; k < e; ++k) {
                        const long long slot = base + (long long)(k - b) * TW + v;
                        const int seg = (int)(std::lower_bound(rows.begin(), rows.end(), s->k_row_idx[k]) - rows.begin());
                        L->sval[(size_t)slot] = s->k_val[k];
                        L->sseg[(size_t)slot] = (uint16_t)seg;
                        // slots ascend in segment: close every cut this slot passes
                        for (; q < ncut && seg >= (q + 1) * TileLayout::CUT; ++q)
                            L->vox_cut[(size_t)((cb + q) * TW + v)] = (uint16_t)(k - b);
                    }
                    for (; q < ncut; ++q) L->vox_cut[(size_t)((cb + q) * TW + v)] = (uint16_t)(e - b);
                }
                std::vector<int>().swap(rows_of[(size_t)it]);
            }
        }
    });
    return true;
}

bool tile_layout_apply(const TileLayout &L, int n, const double *u, double *P) {
    const int S = L.S, TW = L.TW;
    std::atomic<bool> ok{true};
    par_for(L.T, [&](long long tb, long long te) {
        std::vector<double> g;
        for (long long t = tb; t < te; ++t) {
            const int c0 = L.tile_start[(size_t)t], tw = L.tile_start[(size_t)t + 1] - c0;
            for (int sl = 0; sl < S; ++sl) {
                const long long it = t * S + sl;
                const long long s0 = L.seg_start[(size_t)it], ns = L.seg_start[(size_t)it + 1] - s0;
                const long long cb = L.cut_base[(size_t)it], ncut = L.cut_base[(size_t)it + 1] - cb;
                g.resize((size_t)ns);
                for (long long i = 0; i < ns; ++i) g[(size_t)i] = u[L.seg_ray[(size_t)(s0 + i)]];
                // p: voxel (padded form) or position (compact form)
                for (int p = 0; p < tw; ++p) {
                    double acc = 0.0;
                    const int len = L.vox_len[(size_t)it * TW + p];
                    const int v = L.compact ? L.pos_vox[(size_t)it * TW + p] : p;
                    if (v >= tw) ok.store(false);
                    int j = 0;
                    for (long long q = 0; q <= ncut; ++q) {
                        const long long lo = q * TileLayout::CUT, hi = (q + 1) * TileLayout::CUT;
                        const int jend = q < ncut ? L.vox_cut[(size_t)((cb + q) * TW + p)] : len;
                        if (jend < j || jend > len) ok.store(false);
                        for (; j < jend && j < len; ++j) {
                            const long long slot =
                                L.slot_base[(size_t)it] +
                                (L.packed ? (long long)L.row_off[(size_t)(L.row_base[(size_t)it] + j)] : (long long)j * TW) + p;
                            const int sg = L.sseg[(size_t)slot];
                            if (sg < lo || (q < ncut && sg >= hi)) ok.store(false);
                            acc += L.sval[(size_t)slot] * g[(size_t)sg];
                        }
                    }
                    if (j != len) ok.store(false);
                    P[(size_t)sl * n + c0 + (v < tw ? v : 0)] = acc;
                }
            }
        }
    });
    return ok.load();
}

void compact_tile_layout(TileLayout *L, bool pack) {
    if (L->compact) return;
    const int TW = L->TW;
    const long long items = (long long)L->T * L->S;
    std::vector<uint8_t> pos_vox((size_t)items * TW, 0);
    std::vector<uint16_t> plen((size_t)items * TW, 0), pcut(L->vox_cut.size(), 0);
    std::vector<long long> nrow((size_t)items + 1, 0), nslot((size_t)items + 1, 0);
    // positions: voxels by descending length (ties: ascending voxel)
    par_for(items, [&](long long ib, long long ie) {
        std::vector<int> ord(TW);
        for (long long it = ib; it < ie; ++it) {
            const uint16_t *len = L->vox_len.data() + (size_t)it * TW;
            for (int v = 0; v < TW; ++v) ord[(size_t)v] = v;
            std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return len[a] > len[b]; });
            long long tot = 0;
            for (int p = 0; p < TW; ++p) {
                const int v = ord[(size_t)p];
                pos_vox[(size_t)it * TW + p] = (uint8_t)v;
                plen[(size_t)it * TW + p] = len[v];
                tot += len[v];
            }
            for (long long q = L->cut_base[(size_t)it]; q < L->cut_base[(size_t)it + 1]; ++q)
                for (int p = 0; p < TW; ++p) pcut[(size_t)(q * TW + p)] = L->vox_cut[(size_t)(q * TW + ord[(size_t)p])];
            nrow[(size_t)it + 1] = pack ? len[ord[0]] : 0;
            nslot[(size_t)it + 1] = pack ? tot : (long long)len[ord[0]] * TW;
        }
    });
    for (long long it = 0; it < items; ++it) {
        nrow[(size_t)it + 1] += nrow[(size_t)it];
        nslot[(size_t)it + 1] += nslot[(size_t)it];
    }
    std::vector<int> row_off((size_t)nrow[(size_t)items], 0);
    std::vector<double> sval((size_t)nslot[(size_t)items], 0.0);
    std::vector<uint16_t> sseg((size_t)nslot[(size_t)items], 0);
    par_for(items, [&](long long ib, long long ie) {
        for (long long it = ib; it < ie; ++it) {
            const uint16_t *len = plen.data() + (size_t)it * TW;
            const int rows = (int)(nrow[(size_t)it + 1] - nrow[(size_t)it]);
            int *ro = row_off.data() + nrow[(size_t)it];
            // row j holds positions 0 .. c_j - 1, c_j = #{p : len[p] > j}
            int c = TW, off = 0;
            for (int j = 0; j < rows; ++j) {
                while (c > 0 && len[c - 1] <= j) --c;
                ro[j] = off;
                off += c;
            }
            const long long ob = L->slot_base[(size_t)it], nb = nslot[(size_t)it];
            for (int p = 0; p < TW; ++p) {
                const int v = pos_vox[(size_t)it * TW + p];
                for (int j = 0; j < len[p]; ++j) {
                    const long long src = ob + (long long)j * TW + v, dst = nb + (pack ? ro[j] : (long long)j * TW) + p;
                    sval[(size_t)dst] = L->sval[(size_t)src];
                    sseg[(size_t)dst] = L->sseg[(size_t)src];
                }
            }
        }
    });
    L->pos_vox.swap(pos_vox);
    L->vox_len.swap(plen);
    L->vox_cut.swap(pcut);
    L->row_base.swap(nrow);
    L->row_off.swap(row_off);
    L->slot_base.swap(nslot);
    L->sval.swap(sval);
    L->sseg.swap(sseg);
    L->n_slot = L->slot_base[(size_t)items];
    L->compact = true;
    L->packed = pack;
}

bool tile_layout_self_check(const plsqr_system *s, const TileLayout &L, std::string *err) {
    const int n = s->n;
    const std::vector<int> &hb = L.slice_bounds;
    std::vector<double> u((size_t)s->m_k), Pt((size_t)L.S * n), Pr((size_t)L.S * n, 0.0);
    for (int i = 0; i < s->m_k; ++i) u[(size_t)i] = std::sin(0.37 * i + 1.0);
    if (!tile_layout_apply(L, n, u.data(), Pt.data())) {
        *err = "tiled layout self-check failed: round cuts disagree with the slots";
        return false;
    }
    for (int o = 0; o < n; ++o) {
        int sl = 0;
        for (int e = s->k_col_ptr[o]; e < s->k_col_ptr[o + 1]; ++e) {
            const int r = s->k_row_idx[e];
            while (r >= hb[(size_t)sl + 1]) ++sl;
            Pr[(size_t)sl * n + L.perm[(size_t)o]] += s->k_val[e] * u[(size_t)r];
        }
    }
    double dmax = 0.0, nrm = 0.0;
    for (size_t i = 0; i < Pt.size(); ++i) {
        dmax = std::max(dmax, std::fabs(Pt[i] - Pr[i]));
        nrm = std::max(nrm, std::fabs(Pr[i]));
    }
    if (!(dmax <= 1e-12 * (nrm + 1.0))) {
        *err = "tiled layout self-check failed: max diff " + std::to_string(dmax);
        return false;
    }
    return true;
}

bool row_tiles_self_check(const plsqr_system *s, const TileLayout &L, const RowTiles &R, std::string *err) {
    const int n = s->n, mk = s->m_k;
    std::vector<double> x((size_t)n), yt((size_t)mk, 0.0), yr((size_t)mk, 0.0);
    for (int j = 0; j < n; ++j) x[(size_t)j] = std::cos(0.11 * j + 0.5);
    row_tiles_apply(L, R, mk, x.data(), yt.data());
    for (int o = 0; o < n; ++o)
        for (int e = s->k_col_ptr[o]; e < s->k_col_ptr[o + 1]; ++e)
            yr[(size_t)s->k_row_idx[e]] += s->k_val[e] * x[(size_t)L.perm[(size_t)o]];
    double dmax = 0.0, nrm = 0.0;
    for (int r = 0; r < mk; ++r) {
        dmax = std::max(dmax, std::fabs(yt[(size_t)r] - yr[(size_t)r]));
        nrm = std::max(nrm, std::fabs(yr[(size_t)r]));
    }
    if (!(dmax <= 1e-12 * (nrm + 1.0))) {
        *err = "mode-1 tiled layout self-check failed: max diff " + std::to_string(dmax);
        return false;
    }
    return true;
}

void order_tile_columns(const plsqr_system *s, const std::vector<int> &tile_start, std::vector<int> &perm) {
    const int n = s->n;
    std::vector<int> inv((size_t)n);
    for (int o = 0; o < n; ++o) inv[(size_t)perm[(size_t)o]] = o;
    std::vector<int> cols;
    for (size_t t = 0; t + 1 < tile_start.size(); ++t) {
        const int c0 = tile_start[t], c1 = tile_start[t + 1];
        cols.assign(inv.begin() + c0, inv.begin() + c1);
        std::stable_sort(cols.begin(), cols.end(), [&](int a, int b) {
            return s->k_col_ptr[a + 1] - s->k_col_ptr[a] > s->k_col_ptr[b + 1] - s->k_col_ptr[b];
        });
        for (int i = 0; i < c1 - c0; ++i) perm[(size_t)cols[(size_t)i]] = c0 + i;
    }
}

bool build_row_tiles(const TileLayout &L, int m_k, RowTiles *R, std::string *err, int window) {
    const int TW = L.TW, G = RowTiles::G;
    const long long items = (long long)L.T * L.S;
    if (L.compact) {
        *err = "the mode-1 layout is built from the padded slots";
        return false;
    }
    if (TW > 1024) {
        *err = "mode-1 tiles hold at most 1024 voxels";
        return false;
    }
    // per segment: the number of its entries below voxel 256 (j + 1), j < NC
    const int NC = (TW - 1) / 256;
    const int CB = NC == 0 ? 0 : NC == 1 ? 1 : 4;
    R->cut_bytes = CB;
    R->grp_base.assign((size_t)items + 1, 0);
    long long gtot = 0;
    for (long long it = 0; it < items; ++it) {
        const long long ns = L.seg_start[(size_t)it + 1] - L.seg_start[(size_t)it];
        gtot += (ns + G - 1) / G;
        if (gtot > INT_MAX / G) {
            *err = "too many segment groups";
            return false;
        }
        R->grp_base[(size_t)it + 1] = (int)gtot;
    }
    R->n_grp = (int)gtot;
    R->ent_base.assign((size_t)gtot + 1, 0);
    R->seg_len.assign((size_t)gtot * G, 0);
    R->seg_lo.assign((size_t)gtot * G * CB, 255);
    // entries of an item = its slots' nonzeros
    std::vector<long long> item_ent((size_t)items + 1, 0);
    for (long long it = 0; it < items; ++it) {
        long long e = 0;
        for (int v = 0; v < TW; ++v) e += L.vox_len[(size_t)it * TW + v];
        item_ent[(size_t)it + 1] = item_ent[(size_t)it] + e;
    }
    R->n_ent = item_ent[(size_t)items];
    R->ent_vox.assign((size_t)R->n_ent, 0);
    R->ent_val.assign((size_t)R->n_ent, 0.0);
    // Lane order of every item's segments: the segments (ray order) cut into
    // windows of `window`, each window sorted by descending length (stable),
    // so that the 64 segments of a wavefront's group have similar lengths and
    // its lanes finish together.  ord[j] is the item-relative ray-order index
    // of the segment at position j; with sorted windows seg_dst sends its
    // partial back to that index, so segv stays in ray order either way.
    const bool sorted = window > 1;
    if (sorted) R->seg_dst.assign((size_t)gtot * G, 0);
    else R->seg_dst.clear();
    std::atomic<bool> ok{true};
    par_for(items, [&](long long ib, long long ie) {
        std::vector<int> len, ord, lo, pit;
        std::vector<long long> off;
        for (long long it = ib; it < ie; ++it) {
            const int ns = (int)(L.seg_start[(size_t)it + 1] - L.seg_start[(size_t)it]);
            len.assign((size_t)ns, 0);
            lo.assign((size_t)ns * (NC > 0 ? NC : 1), 0);
            const long long base = L.slot_base[(size_t)it];
            for (int v = 0; v < TW; ++v)
                for (int j = 0; j < L.vox_len[(size_t)it * TW + v]; ++j) {
                    const int sg = L.sseg[(size_t)(base + (long long)j * TW + v)];
                    ++len[(size_t)sg];
                    for (int c = v / 256; c < NC; ++c) ++lo[(size_t)sg * NC + c];
                }
            ord.resize((size_t)ns);
            for (int i = 0; i < ns; ++i) ord[(size_t)i] = i;
            if (window > 1)
                for (int w0 = 0; w0 < ns; w0 += window)
                    std::stable_sort(ord.begin() + w0, ord.begin() + std::min(ns, w0 + window),
                                     [&](int a, int b) { return len[(size_t)a] > len[(size_t)b]; });
            pit.resize((size_t)ns);
            for (int j = 0; j < ns; ++j) pit[(size_t)ord[(size_t)j]] = j;
            off.assign((size_t)ns + 1, item_ent[(size_t)it]);  // by position
            for (int j = 0; j < ns; ++j) {
                const int ln = len[(size_t)ord[(size_t)j]];
                if (ln > 255) ok.store(false);
                off[(size_t)j + 1] = off[(size_t)j] + ln;
            }
            const int g0 = R->grp_base[(size_t)it];
            for (int g = 0; g * G < ns; ++g) {
                R->ent_base[(size_t)g0 + g] = off[(size_t)g * G];
                for (int l = 0; l < G && g * G + l < ns; ++l) {
                    const int sg = ord[(size_t)(g * G + l)];
                    R->seg_len[(size_t)(g0 + g) * G + l] = (uint8_t)std::min(len[(size_t)sg], 255);
                    if (sorted) R->seg_dst[(size_t)(g0 + g) * G + l] = (uint16_t)sg;
                    for (int c = 0; c < NC; ++c)
                        R->seg_lo[((size_t)(g0 + g) * G + l) * CB + c] = (uint8_t)std::min(lo[(size_t)sg * NC + c], 255);
                }
            }
            // voxels ascending: each segment's list comes out in voxel order
            for (int v = 0; v < TW; ++v)
                for (int j = 0; j < L.vox_len[(size_t)it * TW + v]; ++j) {
                    const long long slot = base + (long long)j * TW + v;
                    const int sg = L.sseg[(size_t)slot];
                    const long long o = off[(size_t)pit[(size_t)sg]]++;
                    R->ent_vox[(size_t)o] = (uint8_t)(v & 255);
                    R->ent_val[(size_t)o] = L.sval[(size_t)slot];
                }
        }
    });
    if (!ok.load()) {
        *err = "a segment holds more than 255 entries";
        return false;
    }
    R->ent_base[(size_t)gtot] = R->n_ent;
    R->ray_ptr.assign((size_t)m_k + 1, 0);
    for (long long k = 0; k < L.n_seg; ++k) ++R->ray_ptr[(size_t)L.seg_ray[(size_t)k] + 1];
    for (int r = 0; r < m_k; ++r) R->ray_ptr[(size_t)r + 1] += R->ray_ptr[(size_t)r];
    R->ray_seg.assign((size_t)L.n_seg, 0);
    std::vector<int> next(R->ray_ptr.begin(), R->ray_ptr.end() - 1);
    // per ray, its segments in tile order (items ascending), each named by its
    // mode-1 partial's index (the segment's own index: segv is in ray order)
    for (long long k = 0; k < L.n_seg; ++k) R->ray_seg[(size_t)next[(size_t)L.seg_ray[(size_t)k]]++] = (int)k;
    return true;
}

void row_tiles_apply(const TileLayout &L, const RowTiles &R, int m_k, const double *x, double *y) {
    const int G = RowTiles::G;
    const int NC = R.cut_bytes == 0 ? 0 : (L.TW - 1) / 256;
    const long long items = (long long)L.T * L.S;
    std::vector<double> segv((size_t)L.n_seg, 0.0);
    par_for(items, [&](long long ib, long long ie) {
        for (long long it = ib; it < ie; ++it) {
            const int t = (int)(it / L.S);
            const int c0 = L.tile_start[(size_t)t];
            for (int g = R.grp_base[(size_t)it]; g < R.grp_base[(size_t)it + 1]; ++g)
                for (int l = 0; l < G; ++l) {
                    const int len = R.seg_len[(size_t)g * G + l];
                    if (len == 0) continue;
                    long long lo = 0;  // the lengths of the lanes before l
                    for (int q = 0; q < l; ++q) lo += R.seg_len[(size_t)g * G + q];
                    double acc = 0.0;
                    for (int k = 0; k < len; ++k) {
                        const long long o = R.ent_base[(size_t)g] + lo + k;
                        int v = R.ent_vox[(size_t)o];
                        for (int c = 0; c < NC; ++c)
                            if (k >= R.seg_lo[((size_t)g * G + l) * R.cut_bytes + c]) v += 256;
                        acc += R.ent_val[(size_t)o] * x[c0 + v];
                    }
                    const long long dst = R.seg_dst.empty() ? (long long)(g - R.grp_base[(size_t)it]) * G + l
                                                            : (long long)R.seg_dst[(size_t)g * G + l];
                    segv[(size_t)(L.seg_start[(size_t)it] + dst)] = acc;
                }
        }
    });
    for (int r = 0; r < m_k; ++r) {
        double acc = 0.0;
        for (int k = R.ray_ptr[(size_t)r]; k < R.ray_ptr[(size_t)r + 1]; ++k) acc += segv[(size_t)R.ray_seg[(size_t)k]];
        y[r] = acc;
    }
}

namespace {

// Morton code of six coordinates of `bits` bits each (bit i of coordinate j
// at position 6 i + j)
uint64_t morton6(const int *v, int bits) {
    uint64_t r = 0;
    for (int i = 0; i < bits; ++i)
        for (int j = 0; j < 6; ++j) r |= (uint64_t)((v[j] >> i) & 1) << (6 * i + j);
    return r;
}

// Per kernel row, a 6-D Morton key of its two end voxels on an nx*ny*nz grid
// (columns in the loaded numbering, x fastest).  The ends are the voxels
// with the smallest and the largest coordinate along the axis of the row's
// longest extent (the mean of ties, rounded), ordered by (z, y, x) descending
// so both directions of a ray give one key.  Rows sorted by it put rays with
// nearby ends, which cross the same bricks, next to each other.
std::vector<uint64_t> end_keys(const plsqr_system *s, int nx, int ny, int nz) {
    const int mk = s->m_k, n = s->n;
    // the columns of every kernel row, by a transpose with per-thread row
    // counts over column chunks (no shared counters)
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (const char *e = std::getenv("PLSQR_HOST_THREADS")) nt = (unsigned)std::max(1, std::min(16, atoi(e)));
    const long long chunk = ((long long)n + nt - 1) / nt;
    std::vector<std::vector<long long>> off(nt, std::vector<long long>((size_t)mk, 0));
    auto run = [&](auto &&body) {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; ++t) {
            const long long cb = (long long)t * chunk, ce = std::min<long long>(n, cb + chunk);
            th.emplace_back([&body, t, cb, ce] { body(t, cb, ce); });
        }
        for (auto &x : th) x.join();
    };
    run([&](unsigned t, long long cb, long long ce) {
        std::vector<long long> &o = off[t];
        for (long long c = cb; c < ce; ++c)
            for (int e = s->k_col_ptr[c]; e < s->k_col_ptr[c + 1]; ++e) ++o[(size_t)s->k_row_idx[e]];
    });
    std::vector<long long> rp((size_t)mk + 1, 0);
    for (int r = 0; r < mk; ++r) {
        long long at = rp[(size_t)r];
        for (unsigned t = 0; t < nt; ++t) {
            const long long k = off[t][(size_t)r];
            off[t][(size_t)r] = at;
            at += k;
        }
        rp[(size_t)r + 1] = at;
    }
    std::vector<int> cols((size_t)rp[(size_t)mk]);
    run([&](unsigned t, long long cb, long long ce) {
        std::vector<long long> &o = off[t];
        for (long long c = cb; c < ce; ++c)
            for (int e = s->k_col_ptr[c]; e < s->k_col_ptr[c + 1]; ++e) cols[(size_t)o[(size_t)s->k_row_idx[e]]++] = (int)c;
    });
    std::vector<std::vector<long long>>().swap(off);
    int bits = 1;
    while ((1 << bits) < std::max({nx, ny, nz})) ++bits;
    bits = std::min(bits, 10);
    std::vector<uint64_t> key((size_t)mk, 0);
    par_for(mk, [&](long long rb, long long re) {
        for (long long r = rb; r < re; ++r) {
            int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {-1, -1, -1};
            for (long long k = rp[(size_t)r]; k < rp[(size_t)r + 1]; ++k) {
                const int c = cols[(size_t)k];
                const int p[3] = {c % nx, (c / nx) % ny, c / (nx * ny)};
                for (int a = 0; a < 3; ++a) {
                    lo[a] = std::min(lo[a], p[a]);
                    hi[a] = std::max(hi[a], p[a]);
                }
            }
            if (rp[(size_t)r] == rp[(size_t)r + 1]) continue;  // empty row: key 0
            int ax = 0;
            for (int a = 1; a < 3; ++a)
                if (hi[a] - lo[a] > hi[ax] - lo[ax]) ax = a;
            long long sum[2][3] = {{0, 0, 0}, {0, 0, 0}};
            long long cnt[2] = {0, 0};
            for (long long k = rp[(size_t)r]; k < rp[(size_t)r + 1]; ++k) {
                const int c = cols[(size_t)k];
                const int p[3] = {c % nx, (c / nx) % ny, c / (nx * ny)};
                for (int end = 0; end < 2; ++end) {
                    if (p[ax] != (end == 0 ? lo[ax] : hi[ax])) continue;
                    for (int q = 0; q < 3; ++q) sum[end][q] += p[q];
                    ++cnt[end];
                }
            }
            int v[2][3];
            for (int end = 0; end < 2; ++end)
                for (int q = 0; q < 3; ++q) v[end][q] = (int)std::lround((double)sum[end][q] / (double)cnt[end]);
            const bool swap = std::make_tuple(v[1][2], v[1][1], v[1][0]) > std::make_tuple(v[0][2], v[0][1], v[0][0]);
            const int *f = swap ? v[1] : v[0], *g = swap ? v[0] : v[1];
            const int k6[6] = {f[0], f[1], f[2], g[0], g[1], g[2]};
            key[(size_t)r] = morton6(k6, bits);
        }
    });
    return key;
}

}  // namespace

void sort_rays(const plsqr_system *s, const std::vector<int> &perm, const std::vector<int> &tile_start,
               const std::vector<int> &slice_bounds, const plsqr_grid *grid, RaySorted *o, bool sort_damping) {
    const int n = s->n, mk = s->m_k;
    o->rperm.resize((size_t)mk);
    for (int i = 0; i < mk; ++i) o->rperm[(size_t)i] = i;
    std::vector<int> sb = slice_bounds;
    if (sb.size() < 2) sb = {0, mk};
    if (grid && grid->nx > 0 && grid->ny > 0 && grid->nz > 0 && (long long)grid->nx * grid->ny * grid->nz == n) {
        const std::vector<uint64_t> key = end_keys(s, grid->nx, grid->ny, grid->nz);
        for (size_t q = 0; q + 1 < sb.size(); ++q)
            std::stable_sort(o->rperm.begin() + sb[q], o->rperm.begin() + sb[q + 1],
                             [&](int a, int b) { return key[(size_t)a] < key[(size_t)b]; });
    } else {
        std::vector<int> tile_of((size_t)n, 0);
        for (size_t t = 0; t + 1 < tile_start.size(); ++t)
            for (int c = tile_start[t]; c < tile_start[t + 1]; ++c) tile_of[(size_t)c] = (int)t;
        std::vector<int> lo((size_t)mk, INT_MAX), hi((size_t)mk, -1);
        for (int c = 0; c < n; ++c) {
            const int t = tile_of[(size_t)perm[(size_t)c]];
            for (int e = s->k_col_ptr[c]; e < s->k_col_ptr[c + 1]; ++e) {
                const int r = s->k_row_idx[e];
                lo[(size_t)r] = std::min(lo[(size_t)r], t);
                hi[(size_t)r] = std::max(hi[(size_t)r], t);
            }
        }
        for (size_t q = 0; q + 1 < sb.size(); ++q)
            std::stable_sort(o->rperm.begin() + sb[q], o->rperm.begin() + sb[q + 1], [&](int a, int b) {
                return lo[(size_t)a] != lo[(size_t)b] ? lo[(size_t)a] < lo[(size_t)b] : hi[(size_t)a] < hi[(size_t)b];
            });
    }
    o->rinv.resize((size_t)mk);
    for (int i = 0; i < mk; ++i) o->rinv[(size_t)o->rperm[(size_t)i]] = i;
    o->ri.resize((size_t)s->nnz_k);
    o->kv.resize((size_t)s->nnz_k);
    par_for(n, [&](long long cb, long long ce) {
        std::vector<std::pair<int, double>> tmp;
        for (long long c = cb; c < ce; ++c) {
            const int b = s->k_col_ptr[c], e = s->k_col_ptr[c + 1];
            tmp.clear();
            for (int k = b; k < e; ++k) tmp.emplace_back(o->rinv[(size_t)s->k_row_idx[k]], s->k_val[k]);
            std::sort(tmp.begin(), tmp.end(), [](const std::pair<int, double> &a, const std::pair<int, double> &b2) {
                return a.first < b2.first;
            });
            for (int k = b; k < e; ++k) {
                o->ri[(size_t)k] = tmp[(size_t)(k - b)].first;
                o->kv[(size_t)k] = tmp[(size_t)(k - b)].second;
            }
        }
    });
    o->sys = *s;
    o->sys.owned = 0;
    o->sys.k_row_idx = o->ri.data();
    o->sys.k_val = o->kv.data();
    if (!sort_damping) return;
    // damping rows: stable sort by the new column of each row's first entry
    const int md = s->m_d;
    std::vector<int> dperm((size_t)md);
    for (int i = 0; i < md; ++i) dperm[(size_t)i] = i;
    if (md > 0 && (int)perm.size() == n) {
        std::vector<int> key((size_t)md, INT_MAX);
        for (int r = 0; r < md; ++r)
            if (s->d_row_ptr[r] < s->d_row_ptr[r + 1]) key[(size_t)r] = perm[(size_t)s->d_col_idx[s->d_row_ptr[r]]];
        std::stable_sort(dperm.begin(), dperm.end(), [&](int a, int b) { return key[(size_t)a] < key[(size_t)b]; });
    }
    std::vector<int> dinv((size_t)md);
    for (int i = 0; i < md; ++i) dinv[(size_t)dperm[(size_t)i]] = i;
    const long long nd = s->nnz_d;
    o->drp.assign((size_t)md + 1, 0);
    o->dci.resize((size_t)std::max(nd, 1LL));
    o->dv.resize((size_t)std::max(nd, 1LL));
    for (int i = 0; i < md; ++i) {
        const int r = dperm[(size_t)i];
        const int b = s->d_row_ptr[r], e = s->d_row_ptr[r + 1];
        o->drp[(size_t)i + 1] = o->drp[(size_t)i] + (e - b);
        std::copy(s->d_col_idx + b, s->d_col_idx + e, o->dci.begin() + o->drp[(size_t)i]);
        std::copy(s->d_val + b, s->d_val + e, o->dv.begin() + o->drp[(size_t)i]);
    }
    // damping CSC: the same columns, rows renamed and ascending again
    o->dcr.resize((size_t)std::max(nd, 1LL));
    o->dcv.resize((size_t)std::max(nd, 1LL));
    par_for(n, [&](long long cb, long long ce) {
        std::vector<std::pair<int, double>> tmp;
        for (long long c = cb; c < ce; ++c) {
            const int b = s->dc_col_ptr[c], e = s->dc_col_ptr[c + 1];
            tmp.clear();
            for (int k = b; k < e; ++k) tmp.emplace_back(dinv[(size_t)s->dc_row_idx[k]], s->dc_val[k]);
            std::sort(tmp.begin(), tmp.end(), [](const std::pair<int, double> &a, const std::pair<int, double> &b2) {
                return a.first < b2.first;
            });
            for (int k = b; k < e; ++k) {
                o->dcr[(size_t)k] = tmp[(size_t)(k - b)].first;
                o->dcv[(size_t)k] = tmp[(size_t)(k - b)].second;
            }
        }
    });
    o->sys.d_row_ptr = o->drp.data();
    o->sys.d_col_idx = o->dci.data();
    o->sys.d_val = o->dv.data();
    o->sys.dc_row_idx = o->dcr.data();
    o->sys.dc_val = o->dcv.data();
    o->rperm.resize((size_t)mk + md);
    o->rinv.resize((size_t)mk + md);
    for (int i = 0; i < md; ++i) {
        o->rperm[(size_t)mk + i] = mk + dperm[(size_t)i];
        o->rinv[(size_t)mk + i] = mk + dinv[(size_t)i];
    }
}


}  // namespace plsqr
