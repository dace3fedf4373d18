// Builds the GPU factorization schedule: levels, front pool offsets with reuse,
// per-column assembly lists and the batched tile task lists.
#include "chol_plan.hpp"

#include <algorithm>
#include <iterator>
#include <map>
#include <stdexcept>

namespace wrapd {

namespace {

// The FINAL launch's task order: longest first (k range ns for SCHUR, ns - 128 c for
// GBOT), so that the launch's tail is made of short tasks.  (Grouping tasks that share
// operand blocks onto one XCD's L2 measured slower on the C5 beam: the kernel is not
// bound by L2 misses, and the groups unbalance the in-order round-robin dispatch.)
template <typename Push>
void final_order(const std::vector<int>& L, const std::vector<int>& ns, const std::vector<int>& f, Push push) {
    constexpr int T = kFinalTile;
    std::vector<std::pair<int, int>> tk;   // (s, y)
    for (int s : L) {
        const int U = final_row_tiles(ns[s], f[s]);
        const int C = final_col_tiles(ns[s]);
        for (int u = 0; u < U; ++u)
            for (int v = 0; v <= u; ++v) tk.push_back({s, (u << kTaskTileShift) | v});
        for (int u = 0; u < U; ++u)
            for (int c = 0; c < C; ++c) tk.push_back({s, (1 << kTaskKindShift) | (u << kTaskTileShift) | c});
    }
    auto cost = [&](const std::pair<int, int>& x) {
        const int kind = task_kind(x.second);
        return kind == 0 ? ns[x.first] : ns[x.first] - T * (x.second & kTaskTileMask);
    };
    std::stable_sort(tk.begin(), tk.end(),
                     [&](const std::pair<int, int>& a, const std::pair<int, int>& b) { return cost(a) > cost(b); });
    for (const auto& t : tk) push(t.first, t.second);
}

// first-fit allocator over a growing pool, in doubles, with coalescing of free blocks
struct PoolAlloc {
    std::map<int64_t, int64_t> free_blocks;   // offset -> size
    int64_t end = 0;
    static int64_t round(int64_t n) { return (n + 31) / 32 * 32; }
    int64_t alloc(int64_t n) {
        n = round(std::max<int64_t>(n, 1));
        for (auto it = free_blocks.begin(); it != free_blocks.end(); ++it) {
            if (it->second >= n) {
                const int64_t off = it->first;
                const int64_t rest = it->second - n;
                free_blocks.erase(it);
                if (rest > 0) free_blocks[off + n] = rest;
                return off;
            }
        }
        // extend the pool, merging a free tail block if there is one
        if (!free_blocks.empty()) {
            auto last = std::prev(free_blocks.end());
            if (last->first + last->second == end) {
                const int64_t off = last->first;
                free_blocks.erase(last);
                end = off + n;
                return off;
            }
        }
        const int64_t off = end;
        end += n;
        return off;
    }
    void release(int64_t off, int64_t n) {
        n = round(std::max<int64_t>(n, 1));
        auto it = free_blocks.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_blocks.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_blocks.erase(nx);
        }
        if (it != free_blocks.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_blocks.erase(it);
            }
        }
    }
};

}  // namespace

void chol_plan(const CholSymbolic& S, FactorPlan& P, const PlanOptions& opt) {
    const int asm_chunk = opt.asm_chunk;
    const int nsup = S.nsuper;
    P = FactorPlan();
    std::vector<int> ns(nsup), f(nsup);
    for (int s = 0; s < nsup; ++s) {
        ns[s] = S.col0[s + 1] - S.col0[s];
        f[s] = S.fsize[s];
        if (ns[s] <= 0 || f[s] < ns[s]) throw std::runtime_error("chol_plan: bad supernode");
        if (tile_count(ns[s], f[s]) > kTaskTileMask || f[s] >= (1 << 16)) throw std::runtime_error("chol_plan: front too large");
    }
    // levels: leaves 0, parent above every child (parent index > child index)
    std::vector<int> lev(nsup, 0);
    for (int s = 0; s < nsup; ++s) {
        const int p = S.parent[s];
        if (p >= 0) {
            if (p <= s) throw std::runtime_error("chol_plan: supernodes not postordered");
            lev[p] = std::max(lev[p], lev[s] + 1);
        }
    }
    int nlev = 0;
    for (int s = 0; s < nsup; ++s) nlev = std::max(nlev, lev[s] + 1);
    P.n_levels = nlev;
    std::vector<std::vector<int>> by_level(nlev);
    for (int s = 0; s < nsup; ++s) by_level[lev[s]].push_back(s);
    std::vector<std::vector<int>> children(nsup);
    for (int s = 0; s < nsup; ++s)
        if (S.parent[s] >= 0) children[S.parent[s]].push_back(s);

    // pool offsets: a level's fronts are allocated before its assembly; the children it
    // consumes are released after it, for later levels to reuse
    PoolAlloc pool;
    P.foff.assign(nsup, 0);
    for (int l = 0; l < nlev; ++l) {
        for (int s : by_level[l]) {
            P.foff[s] = pool.alloc(static_cast<int64_t>(f[s]) * f[s]);
            P.pool_no_reuse += static_cast<int64_t>(f[s]) * f[s];
        }
        for (int s : by_level[l])
            for (int c : children[s]) pool.release(P.foff[c], static_cast<int64_t>(f[c]) * f[c]);
    }
    P.pool_size = pool.end;

    // assembly lists per global front column
    P.cbase.assign(nsup, 0);
    int64_t ncols = 0;
    for (int s = 0; s < nsup; ++s) {
        P.cbase[s] = ncols;
        ncols += f[s];
    }
    std::vector<int64_t> acnt(ncols + 1, 0), ccnt(ncols + 1, 0);
    const int64_t nnzA = static_cast<int64_t>(S.a_sn.size());
    for (int64_t q = 0; q < nnzA; ++q) {
        const int s = S.a_sn[q];
        const int64_t col = S.a_pos[q] / f[s];
        acnt[P.cbase[s] + col + 1]++;
    }
    for (int s = 0; s < nsup; ++s)
        for (int c : children[s]) {
            const int mc = f[c] - ns[c];
            for (int jj = 0; jj < mc; ++jj) ccnt[P.cbase[s] + S.rel[S.relptr[c] + jj] + 1]++;
        }
    for (int64_t i = 0; i < ncols; ++i) {
        acnt[i + 1] += acnt[i];
        ccnt[i + 1] += ccnt[i];
    }
    // the children's update rows by parent row chunk (rel is increasing along a child's rows)
    if (asm_chunk < 64 || asm_chunk > 1024) throw std::runtime_error("chol_plan: bad assembly chunk");
    P.asm_chunk = asm_chunk;
    P.cchunk_ptr.assign(nsup + 1, 0);
    for (int c = 0; c < nsup; ++c) {
        const int p = S.parent[c];
        P.cchunk_ptr[c + 1] = P.cchunk_ptr[c] + (p >= 0 ? (f[p] + asm_chunk - 1) / asm_chunk + 1 : 0);
    }
    P.cchunk.assign(P.cchunk_ptr[nsup], 0);
    for (int c = 0; c < nsup; ++c) {
        const int p = S.parent[c];
        if (p < 0) continue;
        const int mc = f[c] - ns[c];
        const int nq = (f[p] + asm_chunk - 1) / asm_chunk;
        const int* rl = &S.rel[S.relptr[c]];
        int ii = 0;
        for (int q = 0; q <= nq; ++q) {
            while (ii < mc && rl[ii] < static_cast<int64_t>(asm_chunk) * q) ++ii;
            P.cchunk[P.cchunk_ptr[c] + q] = q == nq ? mc : ii;
        }
    }
    P.acol_ptr = acnt;
    P.ccol_ptr = ccnt;
    P.acol.assign(2 * acnt[ncols], 0);
    P.ccol.assign(2 * ccnt[ncols], 0);
    {
        std::vector<int64_t> pos(acnt.begin(), acnt.end() - 1);
        for (int64_t q = 0; q < nnzA; ++q) {
            const int s = S.a_sn[q];
            const int64_t col = S.a_pos[q] / f[s];
            const int64_t row = S.a_pos[q] % f[s];
            if (row < col) throw std::runtime_error("chol_plan: A entry above the diagonal");
            const int64_t k = pos[P.cbase[s] + col]++;
            P.acol[2 * k] = static_cast<int>(row);
            P.acol[2 * k + 1] = static_cast<int>(q);
        }
        std::vector<int64_t> cp(ccnt.begin(), ccnt.end() - 1);
        for (int s = 0; s < nsup; ++s)
            for (int c : children[s]) {   // ascending child index: the summation order
                const int mc = f[c] - ns[c];
                for (int jj = 0; jj < mc; ++jj) {
                    const int64_t k = cp[P.cbase[s] + S.rel[S.relptr[c] + jj]]++;
                    P.ccol[2 * k] = c;
                    P.ccol[2 * k + 1] = jj;
                }
            }
    }

    // launches
    auto begin = [&](int kind, int panel) {
        P.launches.push_back(FactorLaunch{kind, panel, static_cast<int64_t>(P.tasks.size() / 2), 0, 0});
    };
    auto push = [&](int x, int y) {
        P.tasks.push_back(x);
        P.tasks.push_back(y);
        P.launches.back().count++;
    };
    const double T = kFactorTile;
    std::vector<int> fin_launch(nsup, -1);   // the FINAL launch holding a supernode's tasks
    for (int l = 0; l < nlev; ++l) {
        const std::vector<int>& L = by_level[l];
        // The level's fronts, grouped by the last FINAL launch of the previous level that
        // their children there are in (every front of level l has one): a group whose launch
        // is on the second stream and not the level's last waits for that launch only, so
        // its assembly overlaps the later FINAL launches; the last group waits for all.
        std::map<int, std::vector<int>> groups;
        for (int s : L) {
            int key = -1;
            if (opt.asm_groups)
                for (int c : children[s])
                    if (lev[c] == l - 1) key = std::max(key, fin_launch[c]);
            groups[key].push_back(s);
        }
        int gi = 0;
        for (const auto& [key, grp] : groups) {
            begin(FK_ASSEMBLE, 0);
            const bool last = ++gi == static_cast<int>(groups.size());
            if (!last && key >= 0 && P.launches[key].side) P.launches.back().wait = P.launches[key].ord;
            for (int s : grp) {
                for (int j = 0; j < f[s]; ++j)
                    for (int q = j / asm_chunk; q * asm_chunk < f[s]; ++q) push(s, j | (q << 16));
                P.launches.back().rows = std::max(P.launches.back().rows, std::min(f[s], asm_chunk));
            }
        }
        int nfinal = 0;   // the level's FINAL launches so far
        int pmax = 0;
        for (int s : L) pmax = std::max(pmax, tile_count_l11(ns[s]));
        // Where L21 is formed: by the panel steps (right-looking: each panel's TRSM and updates
        // also cover the update rows), or by one L21 launch per panel k on the second stream,
        // queued once the diagonal tile L_kk is factored: each update-row tile t then applies
        // the updates of panels p < k to F[t, k] in order and solves with L_kk, overlapping the
        // main stream's remaining panel steps (which touch L11 only).  Both apply the same
        // products in the same order, so L21 is the same either way.
        double rl_tiles = 0;
        for (int s : L) {
            const double p11 = tile_count_l11(ns[s]);
            rl_tiles += (tile_count(ns[s], f[s]) - p11) * p11 * (p11 + 1) / 2;
        }
        const bool left = rl_tiles > opt.l21_right_max_tiles * pmax;
        auto l21_launch = [&](int k) {   // after L_kk and the L11 tiles (k, p < k) are final
            if (!left) return;
            begin(FK_L21, k);
            P.launches.back().side = true;
            for (int s : L) {
                const int p11 = tile_count_l11(ns[s]);
                if (k >= p11) continue;
                for (int t = p11; t < tile_count(ns[s], f[s]); ++t) push(s, t);
                P.flops += (tile_count(ns[s], f[s]) - p11) * (k + 1) * 2 * T * T * T;
            }
        };
        // the FINAL launch of the supernodes with p panels: after TRSM(p - 1), their last panel
        // step; on the second stream (after their L21 launches) unless it is the level's last
        // and L21 comes from the panel steps
        auto final_launch = [&](int p) {
            std::vector<int> grp;
            for (int s : L)
                if (tile_count_l11(ns[s]) == p) grp.push_back(s);
            if (grp.empty()) return;
            begin(FK_FINAL, 0);
            P.launches.back().side = left || p < pmax;
            P.launches.back().ord = nfinal++;
            for (int s : grp) fin_launch[s] = static_cast<int>(P.launches.size()) - 1;
            final_order(grp, ns, f, [&](int s, int y) { push(s, y); });
            constexpr double T = kFinalTile;
            for (int s : grp) {   // flops issued (the kernel's k-chunks)
                const int U = final_row_tiles(ns[s], f[s]);
                const int C = final_col_tiles(ns[s]);
                const double kch = kFinalChunk * ((ns[s] + kFinalChunk - 1) / kFinalChunk);
                for (int u = 0; u < U; ++u) {
                    P.flops += (u + 1) * 2.0 * T * T * kch;
                    for (int c = 0; c < C; ++c) P.flops += 2.0 * T * T * (kch - T * c);
                }
            }
        };
        for (int k = 0; k < pmax; ++k) {
            if (k == 0) {   // later panels' diagonal tiles are factored by UPDATE(k - 1)'s task on them
                // one wave per supernode of at most 16 columns, a workgroup per wider one
                for (const int kind : {FK_POTRF16, FK_POTRF}) {
                    bool any = false;
                    for (int s : L) {
                        if ((opt.potrf16 && ns[s] <= kPotrf16Cols) != (kind == FK_POTRF16)) continue;
                        if (!any) begin(kind, k);
                        any = true;
                        push(s, 0);
                    }
                }
                l21_launch(0);
            }
            begin(FK_TRSM, k);
            for (int s : L) {
                if (k >= tile_count_l11(ns[s])) continue;
                const int rows = left ? tile_count_l11(ns[s]) : tile_count(ns[s], f[s]);
                for (int t = k + 1; t < rows; ++t) push(s, t);
                for (int c = 0; c < k; ++c) push(s, (1 << kTaskKindShift) | c);   // X[k, k]: from POTRF
            }
            final_launch(k + 1);
            begin(FK_UPDATE, k);
            // first the tasks on the next panels' diagonal tiles (k + 1, k + 1), which go on to
            // factor them: dispatched first, their Cholesky overlaps the other updates
            for (int s : L)
                if (k + 1 < tile_count_l11(ns[s])) {
                    push(s, ((k + 1) << kTaskTileShift) | (k + 1));
                    P.flops += 2 * T * T * T;
                }
            for (int s : L) {
                const int p11 = tile_count_l11(ns[s]);
                if (k >= p11) continue;
                const int rows = left ? p11 : tile_count(ns[s], f[s]);
                for (int b = k + 1; b < p11; ++b)
                    for (int a = b; a < rows; ++a) {
                        if (a == k + 1 && b == k + 1) continue;   // pushed above
                        push(s, (a << kTaskTileShift) | b);
                        P.flops += 2 * T * T * T;
                    }
                for (int i = k + 1; i < p11; ++i)
                    for (int c = 0; c <= k; ++c) {
                        push(s, (1 << kTaskKindShift) | (i << kTaskTileShift) | c);
                        P.flops += 2 * T * T * T;
                    }
            }
            if (k + 1 < pmax) l21_launch(k + 1);
        }
    }
    // drop empty launches (a level whose panel list is empty for a kind)
    std::vector<FactorLaunch> kept;
    for (const FactorLaunch& L : P.launches)
        if (L.count > 0) kept.push_back(L);
    P.launches.swap(kept);
}

}  // namespace wrapd
