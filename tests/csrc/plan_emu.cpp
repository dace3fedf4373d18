// Test helper (not part of the library): runs the GPU factorization schedule built by
// chol_plan sequentially on the CPU, task by task, with the tile geometry the kernels in
// wrapd_amd/csrc/wd_factor.hip use, and records every read and write so that two tasks
// of one launch touching the same entry (one of them writing) are reported.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../wrapd_amd/csrc/chol_plan.hpp"
#include "../../wrapd_amd/csrc/chol_symbolic.hpp"

using namespace wrapd;

namespace {
CholSymbolic S;
FactorPlan PL;
std::vector<int> Ap, Aj;
std::vector<double> pool, G;
std::vector<int64_t> wF, rF, wG, rG;
int64_t stamp = 0, cur_launch = 0, conflicts = 0;

// Launches on the second stream (FactorLaunch::side) run concurrently with the main
// stream's launches issued after them, until an assembly launch waits for them (all of them,
// or those up to the previous level's FINAL launch FactorLaunch::wait): an access of such a
// main launch to an entry a running side launch wrote (or a write to one it read) counts too.
std::vector<char> side_active;   // by launch index + 1
bool cur_main = true;
inline int64_t launch_of(int64_t st) { return st >> 32; }
inline bool concurrent(int64_t st) {
    const int64_t l = launch_of(st);
    return l == cur_launch ? st != stamp : (cur_main && l < static_cast<int64_t>(side_active.size()) && side_active[l]);
}
inline void on_read(std::vector<int64_t>& w, std::vector<int64_t>& r, int64_t e) {
    if (w[e] && concurrent(w[e])) conflicts++;
    r[e] = stamp;
}
inline void on_write(std::vector<int64_t>& w, std::vector<int64_t>& r, int64_t e) {
    if (w[e] && concurrent(w[e])) conflicts++;
    if (r[e] && concurrent(r[e])) conflicts++;
    w[e] = stamp;
}
inline double rdF(int64_t e) { on_read(wF, rF, e); return pool[e]; }
inline void wrF(int64_t e, double v) { on_write(wF, rF, e); pool[e] = v; }
inline double rdG(int64_t e) { on_read(wG, rG, e); return G[e]; }
inline void wrG(int64_t e, double v) { on_write(wG, rG, e); G[e] = v; }

void tile(int t, int ns, int f, int& st, int& sz) {
    const int P = tile_count_l11(ns);
    if (t < P) { st = kFactorTile * t; sz = std::min(kFactorTile, ns - st); }
    else { st = ns + kFactorTile * (t - P); sz = std::min(kFactorTile, f - st); }
}
}  // namespace

PlanOptions popt;
extern "C" void pe_set_options(double l21_right_max_tiles, int asm_chunk, int asm_groups, int potrf16) {
    popt.l21_right_max_tiles = l21_right_max_tiles;
    popt.asm_chunk = asm_chunk;
    popt.asm_groups = asm_groups != 0;
    popt.potrf16 = potrf16 != 0;
}

extern "C" int pe_analyze(int n, const int* ap, const int* aj, const double* xyz, int leaf_size, double* st) {
    Ap.assign(ap, ap + n + 1);
    Aj.assign(aj, aj + ap[n]);
    SymbolicOptions o;
    o.leaf_size = leaf_size;
    chol_analyze(n, Ap, Aj, xyz, o, S);
    chol_plan(S, PL, popt);
    st[0] = S.nsuper; st[1] = PL.n_levels; st[2] = static_cast<double>(PL.launches.size());
    st[3] = S.max_ns; st[4] = static_cast<double>(S.goff[S.nsuper]); st[5] = static_cast<double>(PL.pool_size);
    st[6] = static_cast<double>(PL.pool_no_reuse);
    st[7] = static_cast<double>(std::count_if(PL.launches.begin(), PL.launches.end(),
                                              [](const FactorLaunch& L) { return L.side; }));
    st[8] = static_cast<double>(std::count_if(PL.launches.begin(), PL.launches.end(),
                                              [](const FactorLaunch& L) { return L.kind == FK_ASSEMBLE && L.wait >= 0; }));
    st[9] = static_cast<double>(std::count_if(PL.launches.begin(), PL.launches.end(),
                                              [](const FactorLaunch& L) { return L.kind == FK_L21; }));
    return 0;
}

// perm (n), col0 (nsuper+1), fsize (nsuper), goff (nsuper+1), rowptr (nsuper+1), rows
extern "C" void pe_layout(int* perm, int* col0, int* fsize, int64_t* goff, int64_t* rowptr, int* rows) {
    std::copy(S.perm.begin(), S.perm.end(), perm);
    std::copy(S.col0.begin(), S.col0.end(), col0);
    std::copy(S.fsize.begin(), S.fsize.end(), fsize);
    std::copy(S.goff.begin(), S.goff.end(), goff);
    std::copy(S.rowptr.begin(), S.rowptr.end(), rowptr);
    if (rows) std::copy(S.rows.begin(), S.rows.end(), rows);
}

// returns 0 on success, 1 on a non-positive pivot, 2 on an L21 task outside the update rows,
// 3 on a child row outside its assembly chunk;
// G_out gets G; conflicts_out the count
extern "C" int pe_run(const double* ax, double* G_out, int64_t* conflicts_out) {
    pool.assign(PL.pool_size, std::nan(""));
    G.assign(S.goff[S.nsuper], std::nan(""));
    wF.assign(pool.size(), 0); rF.assign(pool.size(), 0);
    wG.assign(G.size(), 0); rG.assign(G.size(), 0);
    conflicts = 0;
    int fail = 0;
    side_active.assign(PL.launches.size() + 1, 0);
    std::vector<int64_t> fin_prev, fin_cur;   // launch index + 1 of the previous / this level's FINAL launch k
    for (size_t li = 0; li < PL.launches.size(); ++li) {
        const FactorLaunch& L = PL.launches[li];
        cur_launch = static_cast<int64_t>(li) + 1;
        if (L.kind == FK_ASSEMBLE) {
            if (li > 0 && PL.launches[li - 1].kind != FK_ASSEMBLE) {   // a new level
                fin_prev = fin_cur;
                fin_cur.clear();
            }
            // the second stream runs in order: waiting for its launch w waits for all before it
            const int64_t upto = L.wait < 0 ? cur_launch
                                 : L.wait < static_cast<int>(fin_prev.size()) ? fin_prev[L.wait] : -1;
            if (upto < 0 || (L.wait >= 0 && !PL.launches[upto - 1].side)) return 3;   // not a second-stream launch
            std::fill(side_active.begin(), side_active.begin() + upto + 1, 0);
        }
        if (L.kind == FK_FINAL) {
            if (L.ord < 0) return 3;
            if (static_cast<int>(fin_cur.size()) <= L.ord) fin_cur.resize(L.ord + 1, -1);
            fin_cur[L.ord] = cur_launch;
        }
        cur_main = !L.side;
        if (L.side) side_active[cur_launch] = 1;
        for (int q = 0; q < L.count; ++q) {
            stamp = (cur_launch << 32) | (q + 1);
            const int s = PL.tasks[2 * (L.offset + q)], y = PL.tasks[2 * (L.offset + q) + 1];
            const int ns = S.col0[s + 1] - S.col0[s], f = S.fsize[s];
            const int64_t fo = PL.foff[s], go = S.goff[s];
            const int k = L.panel, c0 = kFactorTile * k, nb = std::min(kFactorTile, ns - c0);
            auto Fi = [&](int r, int c) { return fo + static_cast<int64_t>(c) * f + r; };
            auto Gi = [&](int r, int c) { return go + static_cast<int64_t>(c) * g_ld(f) + r; };
            // Cholesky of panel kp's diagonal tile in place, and its inverse into G's tile (kp, kp)
            auto potrf = [&](int kp) {
                const int cp = kFactorTile * kp, nbp = std::min(kFactorTile, ns - cp);
                for (int p = 0; p < nbp; ++p) {
                    double d = rdF(Fi(cp + p, cp + p));
                    if (!(d > 0)) { fail = 1; d = 1.0; }
                    d = std::sqrt(d);
                    wrF(Fi(cp + p, cp + p), d);
                    for (int r = p + 1; r < nbp; ++r) wrF(Fi(cp + r, cp + p), rdF(Fi(cp + r, cp + p)) / d);
                    for (int c = p + 1; c < nbp; ++c)
                        for (int r = c; r < nbp; ++r)
                            wrF(Fi(cp + r, cp + c), rdF(Fi(cp + r, cp + c)) - rdF(Fi(cp + r, cp + p)) * rdF(Fi(cp + c, cp + p)));
                }
                std::vector<double> X(kFactorTile * kFactorTile, 0.0);
                for (int j = 0; j < nbp; ++j) X[j * 64 + j] = 1.0;
                for (int p = 0; p < nbp; ++p) {
                    const double d = rdF(Fi(cp + p, cp + p));
                    for (int j = 0; j < nbp; ++j) X[j * 64 + p] /= d;
                    for (int r = p + 1; r < nbp; ++r) {
                        const double l = rdF(Fi(cp + r, cp + p));
                        for (int j = 0; j < nbp; ++j) X[j * 64 + r] -= l * X[j * 64 + p];
                    }
                }
                for (int j = 0; j < nbp; ++j)
                    for (int r = 0; r < nbp; ++r) wrG(Gi(cp + r, cp + j), X[j * 64 + r]);
            };
            if (L.kind == FK_ASSEMBLE) {   // rows [max(j, q C), min(f, (q + 1) C)) of column j
                const int j = y & 0xFFFF, qc = y >> 16, C = PL.asm_chunk;
                const int r0 = std::max(j, qc * C), r1 = std::min(f, (qc + 1) * C);
                const int64_t gc = PL.cbase[s] + j;
                for (int r = r0; r < r1; ++r) wrF(Fi(r, j), 0.0);
                if (j < ns && qc == j / C) for (int r = 0; r < ns; ++r) wrG(Gi(r, j), 0.0);   // L11^{-1} rows only
                for (int64_t a = PL.acol_ptr[gc]; a < PL.acol_ptr[gc + 1]; ++a)
                    if (PL.acol[2 * a] >= r0 && PL.acol[2 * a] < r1) wrF(Fi(PL.acol[2 * a], j), ax[PL.acol[2 * a + 1]]);
                for (int64_t a = PL.ccol_ptr[gc]; a < PL.ccol_ptr[gc + 1]; ++a) {
                    const int ch = PL.ccol[2 * a], jj = PL.ccol[2 * a + 1];
                    const int fc = S.fsize[ch], nsc = S.col0[ch + 1] - S.col0[ch];
                    const int* rl = &S.rel[S.relptr[ch]];
                    const int* cs = &PL.cchunk[PL.cchunk_ptr[ch]];
                    for (int ii = std::max(jj, cs[qc]); ii < cs[qc + 1]; ++ii) {
                        if (rl[ii] < r0 || rl[ii] >= r1) return 3;   // the chunk table is wrong
                        const double v = rdF(PL.foff[ch] + static_cast<int64_t>(nsc + jj) * fc + nsc + ii);
                        wrF(Fi(rl[ii], j), rdF(Fi(rl[ii], j)) + v);
                    }
                }
            } else if (L.kind == FK_POTRF || L.kind == FK_POTRF16) {
                if (L.kind == FK_POTRF16 && S.col0[s + 1] - S.col0[s] > kPotrf16Cols) return 3;
                potrf(k);
            } else if (L.kind == FK_TRSM) {
                const int kind = task_kind(y), t = y & ((1 << kTaskKindShift) - 1);
                int t0, nt;
                tile(t, ns, f, t0, nt);
                if (kind == 0) {
                    for (int p = 0; p < nb; ++p) {
                        const double d = rdF(Fi(c0 + p, c0 + p));
                        for (int r = 0; r < nt; ++r) wrF(Fi(t0 + r, c0 + p), rdF(Fi(t0 + r, c0 + p)) / d);
                        for (int c = p + 1; c < nb; ++c) {
                            const double l = rdF(Fi(c0 + c, c0 + p));
                            for (int r = 0; r < nt; ++r)
                                wrF(Fi(t0 + r, c0 + c), rdF(Fi(t0 + r, c0 + c)) - rdF(Fi(t0 + r, c0 + p)) * l);
                        }
                    }
                } else if (t != k) {   // the diagonal tile X[k, k] was stored by POTRF
                    std::vector<double> X(kFactorTile * kFactorTile, 0.0);
                    for (int j = 0; j < nt; ++j)
                        for (int r = 0; r < nb; ++r) X[j * 64 + r] = rdG(Gi(c0 + r, t0 + j));
                    for (int p = 0; p < nb; ++p) {
                        const double d = rdF(Fi(c0 + p, c0 + p));
                        for (int j = 0; j < nt; ++j) X[j * 64 + p] /= d;
                        for (int r = p + 1; r < nb; ++r) {
                            const double l = rdF(Fi(c0 + r, c0 + p));
                            for (int j = 0; j < nt; ++j) X[j * 64 + r] -= l * X[j * 64 + p];
                        }
                    }
                    for (int j = 0; j < nt; ++j)
                        for (int r = 0; r < nb; ++r) wrG(Gi(c0 + r, t0 + j), X[j * 64 + r]);
                }
            } else if (L.kind == FK_L21) {   // the 64-row tile t = y >= P at panel k
                const int P = tile_count_l11(ns);
                if (y < P || k >= P) return 2;
                int t0, nt;
                tile(y, ns, f, t0, nt);
                for (int p = 0; p < k; ++p) {   // F[t, k] -= L[t, p] L[k, p]^T
                    const int cp = kFactorTile * p;
                    for (int j = 0; j < nb; ++j)
                        for (int i = 0; i < nt; ++i) {
                            double acc = 0;
                            for (int kk = 0; kk < kFactorTile; ++kk)
                                acc += rdF(Fi(t0 + i, cp + kk)) * rdF(Fi(c0 + j, cp + kk));
                            wrF(Fi(t0 + i, c0 + j), rdF(Fi(t0 + i, c0 + j)) - acc);
                        }
                }
                for (int p = 0; p < nb; ++p) {   // L[t, k] = F[t, k] L_kk^{-T}
                    const double d = rdF(Fi(c0 + p, c0 + p));
                    for (int r = 0; r < nt; ++r) wrF(Fi(t0 + r, c0 + p), rdF(Fi(t0 + r, c0 + p)) / d);
                    for (int c = p + 1; c < nb; ++c) {
                        const double l = rdF(Fi(c0 + c, c0 + p));
                        for (int r = 0; r < nt; ++r)
                            wrF(Fi(t0 + r, c0 + c), rdF(Fi(t0 + r, c0 + c)) - rdF(Fi(t0 + r, c0 + p)) * l);
                    }
                }
            } else if (L.kind == FK_FINAL) {   // kFinalTile x kFinalTile tiles
                const int kind = task_kind(y), T = kFinalTile;
                const int u = (y >> kTaskTileShift) & kTaskTileMask, v = y & kTaskTileMask;
                const int a0 = ns + T * u, na = std::min(T, f - a0);
                const int b0 = kind == 0 ? ns + T * v : T * v;
                const int nbt = std::min(T, (kind == 0 ? f : ns) - b0);
                for (int j = 0; j < nbt; ++j)
                    for (int i = (kind == 0 && u == v) ? j : 0; i < na; ++i) {
                        double acc = 0;
                        for (int q = kind == 0 ? 0 : b0; q < ns; ++q)
                            acc += rdF(Fi(a0 + i, q)) * (kind == 0 ? rdF(Fi(b0 + j, q)) : rdG(Gi(q, b0 + j)));
                        if (kind == 0) wrF(Fi(a0 + i, b0 + j), rdF(Fi(a0 + i, b0 + j)) - acc);
                        else wrG(Gi(a0 + i, b0 + j), -acc);
                    }
            } else {
                const int kind = task_kind(y);
                const int a = (y >> kTaskTileShift) & kTaskTileMask, b = y & kTaskTileMask;
                int a0, na, b0, nbt;
                tile(a, ns, f, a0, na);
                tile(b, ns, f, b0, nbt);
                for (int j = 0; j < nbt; ++j)
                    for (int i = (kind == 0 && a == b) ? j : 0; i < na; ++i) {
                        double acc = 0;
                        for (int kk = 0; kk < nb; ++kk)
                            acc += rdF(Fi(a0 + i, c0 + kk)) * (kind == 0 ? rdF(Fi(b0 + j, c0 + kk)) : rdG(Gi(c0 + kk, b0 + j)));
                        if (kind == 0) wrF(Fi(a0 + i, b0 + j), rdF(Fi(a0 + i, b0 + j)) - acc);
                        else wrG(Gi(a0 + i, b0 + j), rdG(Gi(a0 + i, b0 + j)) - acc);
                    }
                if (kind == 0 && a == b && a == k + 1) potrf(k + 1);   // the next panel's Cholesky
            }
        }
    }
    std::memcpy(G_out, G.data(), G.size() * sizeof(double));
    *conflicts_out = conflicts;
    return fail;
}
