// Host symbolic analysis for the supernodal Cholesky (layout described in the header).
#include "chol_symbolic.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <stdexcept>

namespace wrapd {

namespace {

struct Graph {
    std::vector<int> ptr, adj;   // symmetric, no diagonal
};

Graph symmetric_graph(int n, const std::vector<int>& Ap, const std::vector<int>& Aj) {
    Graph g;
    std::vector<int> deg(n, 0);
    for (int i = 0; i < n; ++i)
        for (int p = Ap[i]; p < Ap[i + 1]; ++p)
            if (Aj[p] != i) { deg[i]++; deg[Aj[p]]++; }
    g.ptr.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) g.ptr[i + 1] = g.ptr[i] + deg[i];
    g.adj.assign(g.ptr[n], 0);
    std::vector<int> fill(g.ptr.begin(), g.ptr.end() - 1);
    for (int i = 0; i < n; ++i)
        for (int p = Ap[i]; p < Ap[i + 1]; ++p) {
            int j = Aj[p];
            if (j == i) continue;
            g.adj[fill[i]++] = j;
            g.adj[fill[j]++] = i;
        }
    return g;
}

// Nested dissection: split the part in two halves, take the smaller one-sided vertex
// separator, recurse, order the separator last.  With vertex coordinates the halves are
// the two sides of the median along the longest bounding-box axis (geometric nested
// dissection); without them, the two sides of the median level of a breadth-first level
// structure rooted at a pseudo-peripheral vertex of the part (George 1973).
struct Dissector {
    const Graph& g;
    const double* xyz;
    int leaf;
    std::vector<int> side;    // -1 outside the current call
    std::vector<int> order;
    std::vector<int> level;   // breadth-first level inside the current part (no coordinates)
    std::vector<int> queue;

    Dissector(const Graph& gg, const double* p, int n, int lf) : g(gg), xyz(p), leaf(lf), side(n, -1) {
        order.reserve(n);
        if (!xyz) level.assign(n, -1);
    }

    // breadth-first levels from `root` over the vertices marked side == 3; returns the last
    // vertex reached (one of the farthest)
    int bfs(int root) {
        queue.clear();
        queue.push_back(root);
        level[root] = 0;
        int last = root;
        for (size_t h = 0; h < queue.size(); ++h) {
            const int v = queue[h];
            last = v;
            for (int p = g.ptr[v]; p < g.ptr[v + 1]; ++p) {
                const int w = g.adj[p];
                if (side[w] == 3 && level[w] < 0) {
                    level[w] = level[v] + 1;
                    queue.push_back(w);
                }
            }
        }
        return last;
    }

    void run(std::vector<int>& ids) {
        if (static_cast<int>(ids.size()) <= leaf) {
            order.insert(order.end(), ids.begin(), ids.end());
            return;
        }
        size_t mid = ids.size() / 2;
        if (xyz) {
            double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
            for (int v : ids)
                for (int c = 0; c < 3; ++c) {
                    lo[c] = std::min(lo[c], xyz[3 * v + c]);
                    hi[c] = std::max(hi[c], xyz[3 * v + c]);
                }
            int ax = 0;
            for (int c = 1; c < 3; ++c) if (hi[c] - lo[c] > hi[ax] - lo[ax]) ax = c;
            std::nth_element(ids.begin(), ids.begin() + mid, ids.end(), [&](int a, int b) {
                double xa = xyz[3 * a + ax], xb = xyz[3 * b + ax];
                return xa < xb || (xa == xb && a < b);
            });
        } else {
            for (int v : ids) { side[v] = 3; level[v] = -1; }
            const int far = bfs(*std::min_element(ids.begin(), ids.end()));
            for (int v : ids) level[v] = -1;
            bfs(far);
            // vertices the search did not reach (another component) go last
            for (int v : ids) if (level[v] < 0) level[v] = std::numeric_limits<int>::max();
            std::nth_element(ids.begin(), ids.begin() + mid, ids.end(), [&](int a, int b) {
                return level[a] < level[b] || (level[a] == level[b] && a < b);
            });
        }
        for (size_t i = 0; i < ids.size(); ++i) side[ids[i]] = i < mid ? 0 : 1;
        std::vector<int> sepL, sepR;
        for (int v : ids) {
            int sv = side[v];
            for (int p = g.ptr[v]; p < g.ptr[v + 1]; ++p) {
                int w = g.adj[p];
                if (side[w] >= 0 && side[w] != sv) {
                    (sv == 0 ? sepL : sepR).push_back(v);
                    break;
                }
            }
        }
        std::vector<int>& sep = sepL.size() <= sepR.size() ? sepL : sepR;
        for (int v : sep) side[v] = 2;
        std::vector<int> L, R;
        L.reserve(mid);
        R.reserve(ids.size() - mid);
        for (int v : ids) {
            if (side[v] == 0) L.push_back(v);
            else if (side[v] == 1) R.push_back(v);
        }
        for (int v : ids) side[v] = -1;
        std::vector<int> S = sep;
        ids.clear();
        ids.shrink_to_fit();
        if (L.empty() || R.empty()) {   // degenerate split: stop here
            order.insert(order.end(), L.begin(), L.end());
            order.insert(order.end(), R.begin(), R.end());
        } else {
            run(L);
            run(R);
        }
        std::sort(S.begin(), S.end(), [&](int a, int b) {
            if (xyz)
                for (int c = 0; c < 3; ++c) {
                    if (xyz[3 * a + c] != xyz[3 * b + c]) return xyz[3 * a + c] < xyz[3 * b + c];
                }
            return a < b;
        });
        order.insert(order.end(), S.begin(), S.end());
    }
};

}  // namespace

void chol_analyze(int n, const std::vector<int>& Ap, const std::vector<int>& Aj, const double* xyz,
                  const SymbolicOptions& opt, CholSymbolic& S) {
    S = CholSymbolic();
    S.n = n;
    Graph g = symmetric_graph(n, Ap, Aj);

    // 1. ordering
    std::vector<int> perm;
    {
        Dissector d(g, xyz, n, opt.leaf_size);
        std::vector<int> all(n);
        std::iota(all.begin(), all.end(), 0);
        d.run(all);
        perm = std::move(d.order);
    }
    std::vector<int> iperm(n);
    for (int k = 0; k < n; ++k) iperm[perm[k]] = k;

    // 2. elimination tree (Liu's algorithm with path compression)
    std::vector<int> parent(n, -1), anc(n, -1);
    for (int i = 0; i < n; ++i) {
        int v = perm[i];
        for (int p = g.ptr[v]; p < g.ptr[v + 1]; ++p) {
            int k = iperm[g.adj[p]];
            if (k >= i) continue;
            int r = k;
            while (anc[r] != -1 && anc[r] != i) {
                int t = anc[r];
                anc[r] = i;
                r = t;
            }
            if (anc[r] == -1) {
                anc[r] = i;
                parent[r] = i;
            }
        }
    }

    // 3. postorder
    std::vector<int> post(n);
    {
        std::vector<int> head(n, -1), next(n, -1);
        for (int j = n - 1; j >= 0; --j)
            if (parent[j] >= 0) { next[j] = head[parent[j]]; head[parent[j]] = j; }
        std::vector<int> stack;
        int k = 0;
        for (int r = 0; r < n; ++r) {
            if (parent[r] != -1) continue;
            stack.push_back(r);
            while (!stack.empty()) {
                int p = stack.back();
                int c = head[p];
                if (c == -1) {
                    stack.pop_back();
                    post[k++] = p;
                } else {
                    head[p] = next[c];
                    stack.push_back(c);
                }
            }
        }
    }
    {
        std::vector<int> ipost(n), perm2(n), parent2(n);
        for (int k = 0; k < n; ++k) ipost[post[k]] = k;
        for (int k = 0; k < n; ++k) {
            perm2[k] = perm[post[k]];
            parent2[k] = parent[post[k]] < 0 ? -1 : ipost[parent[post[k]]];
        }
        perm.swap(perm2);
        parent.swap(parent2);
        for (int k = 0; k < n; ++k) iperm[perm[k]] = k;
    }

    // 4. column counts by row subtrees
    std::vector<int> cnt(n, 1), mark(n, -1), nchild(n, 0);
    for (int j = 0; j < n; ++j) if (parent[j] >= 0) nchild[parent[j]]++;
    for (int i = 0; i < n; ++i) {
        mark[i] = i;
        int v = perm[i];
        for (int p = g.ptr[v]; p < g.ptr[v + 1]; ++p) {
            int j = iperm[g.adj[p]];
            if (j >= i) continue;
            while (mark[j] != i) {
                cnt[j]++;
                mark[j] = i;
                j = parent[j];
            }
        }
    }

    // 5. fundamental supernodes
    std::vector<int> first, nsv, mv;
    for (int j = 0; j < n; ++j) {
        bool merge = j > 0 && parent[j - 1] == j && cnt[j - 1] == cnt[j] + 1 && nchild[j] == 1;
        if (!merge) {
            first.push_back(j);
            nsv.push_back(1);
        } else {
            nsv.back()++;
        }
    }
    int nf = static_cast<int>(first.size());
    for (int s = 0; s < nf; ++s) mv.push_back(cnt[first[s]] - nsv[s]);
    std::vector<int> sn_of(n);
    for (int s = 0; s < nf; ++s)
        for (int j = first[s]; j < first[s] + nsv[s]; ++j) sn_of[j] = s;
    std::vector<int> spar(nf, -1);
    for (int s = 0; s < nf; ++s) {
        int last = first[s] + nsv[s] - 1;
        spar[s] = parent[last] < 0 ? -1 : sn_of[parent[last]];
    }

    // 6. relaxed amalgamation (merge the contiguous last child into its parent)
    std::vector<char> alive(nf, 1);
    std::vector<int> ends_at(n + 1, -1);
    for (int s = 0; s < nf; ++s) ends_at[first[s] + nsv[s]] = s;
    std::vector<int> rep(nf);
    std::iota(rep.begin(), rep.end(), 0);
    auto find = [&](int s) {
        while (rep[s] != s) { rep[s] = rep[rep[s]]; s = rep[s]; }
        return s;
    };
    for (int p = 0; p < nf; ++p) {
        while (true) {
            int c = ends_at[first[p]];
            if (c < 0 || !alive[c] || c == p) break;
            if (find(spar[c] < 0 ? c : spar[c]) != p) break;
            long long z = static_cast<long long>(nsv[c]) * (nsv[p] + mv[p] - mv[c]);
            int w = nsv[c] + nsv[p];
            double tot = 0.5 * w * (w + 1.0) + static_cast<double>(w) * mv[p];
            bool ok = (w <= opt.relax_small) || (w <= opt.relax_max && z <= opt.relax_frac * tot);
            if (!ok) break;
            ends_at[first[p]] = -1;
            first[p] = first[c];
            nsv[p] = w;
            alive[c] = 0;
            rep[c] = p;
        }
    }
    // final supernode columns (alive ones, in column order), split wide ones
    std::vector<int> c0;
    for (int s = 0; s < nf; ++s) {
        if (!alive[s]) continue;
        for (int j = first[s]; j < first[s] + nsv[s]; j += opt.ns_max) c0.push_back(j);
    }
    std::sort(c0.begin(), c0.end());
    c0.push_back(n);
    int ns_total = static_cast<int>(c0.size()) - 1;
    S.nsuper = ns_total;
    S.col0 = c0;
    for (int s = 0; s < ns_total; ++s)
        for (int j = c0[s]; j < c0[s + 1]; ++j) sn_of[j] = s;

    // 7. row structures, supernodal parent
    S.rowptr.assign(ns_total + 1, 0);
    S.parent.assign(ns_total, -1);
    std::vector<std::vector<int>> children(ns_total);
    std::vector<int> mk(n, -1);
    std::vector<int> R;
    std::vector<std::vector<int>> srows(ns_total);
    for (int s = 0; s < ns_total; ++s) {
        int a = c0[s], b = c0[s + 1];
        R.clear();
        for (int j = a; j < b; ++j) {
            int v = perm[j];
            for (int p = g.ptr[v]; p < g.ptr[v + 1]; ++p) {
                int i = iperm[g.adj[p]];
                if (i >= b && mk[i] != s) { mk[i] = s; R.push_back(i); }
            }
        }
        for (int c : children[s]) {
            const std::vector<int>& rc = srows[c];
            int nsc = c0[c + 1] - c0[c];
            for (size_t q = nsc; q < rc.size(); ++q) {
                int i = rc[q];
                if (i >= b && mk[i] != s) { mk[i] = s; R.push_back(i); }
            }
        }
        std::sort(R.begin(), R.end());
        std::vector<int>& rows = srows[s];
        rows.resize(b - a + R.size());
        for (int j = a; j < b; ++j) rows[j - a] = j;
        std::copy(R.begin(), R.end(), rows.begin() + (b - a));
        if (!R.empty()) {
            int p = sn_of[R[0]];
            S.parent[s] = p;
            children[p].push_back(s);
        }
        // children's structures are no longer needed once the parent is built
    }
    S.fsize.resize(ns_total);
    for (int s = 0; s < ns_total; ++s) {
        S.fsize[s] = static_cast<int>(srows[s].size());
        S.rowptr[s + 1] = S.rowptr[s] + S.fsize[s];
    }
    S.rows.resize(S.rowptr[ns_total]);
    for (int s = 0; s < ns_total; ++s)
        std::copy(srows[s].begin(), srows[s].end(), S.rows.begin() + S.rowptr[s]);
    srows.clear();

    // 8. offsets, statistics
    S.goff.assign(ns_total + 1, 0);
    S.uoff.assign(ns_total + 1, 0);
    S.relptr.assign(ns_total + 1, 0);
    for (int s = 0; s < ns_total; ++s) {
        int ns = c0[s + 1] - c0[s];
        int f = S.fsize[s];
        int m = f - ns;
        S.goff[s + 1] = S.goff[s] + static_cast<int64_t>(g_ld(f)) * ns;
        S.uoff[s + 1] = S.uoff[s] + m;
        S.relptr[s + 1] = S.relptr[s] + m;
        S.max_front = std::max(S.max_front, f);
        S.max_ns = std::max(S.max_ns, ns);
        S.nnz_L += 0.5 * ns * (ns + 1.0) + static_cast<double>(ns) * m;
        S.flops_factor += ns * (double)ns * ns / 3.0 + (double)ns * ns * m + (double)ns * m * m;
        S.flops_g += ns * (double)ns * ns / 3.0 + (double)m * ns * ns;
        S.g_entries += static_cast<double>(f) * ns;
    }
    // 9. extend-add maps and forward assembly lists
    S.rel.assign(S.relptr[ns_total], 0);
    std::vector<int64_t> acnt(S.rows.size() + 1, 0);
    for (int c = 0; c < ns_total; ++c) {
        int p = S.parent[c];
        if (p < 0) continue;
        int nsc = c0[c + 1] - c0[c];
        const int* pr = &S.rows[S.rowptr[p]];
        int fp = S.fsize[p];
        for (int q = nsc; q < S.fsize[c]; ++q) {
            int row = S.rows[S.rowptr[c] + q];
            int pos = static_cast<int>(std::lower_bound(pr, pr + fp, row) - pr);
            if (pos >= fp || pr[pos] != row) throw std::runtime_error("chol_analyze: structure not nested");
            S.rel[S.relptr[c] + (q - nsc)] = pos;
            acnt[S.rowptr[p] + pos + 1]++;
        }
    }
    for (size_t i = 0; i < S.rows.size(); ++i) acnt[i + 1] += acnt[i];
    S.asm_ptr = acnt;
    S.asm_idx.assign(acnt.back(), 0);
    {
        std::vector<int64_t> fill(acnt.begin(), acnt.end() - 1);
        for (int c = 0; c < ns_total; ++c) {   // children in increasing order: fixed summation order
            int p = S.parent[c];
            if (p < 0) continue;
            int nsc = c0[c + 1] - c0[c];
            for (int q = nsc; q < S.fsize[c]; ++q) {
                int pos = S.rel[S.relptr[c] + (q - nsc)];
                S.asm_idx[fill[S.rowptr[p] + pos]++] = S.uoff[c] + (q - nsc);
            }
        }
    }
    // 10. A slot -> front position
    int nnzA = Ap[n];
    S.a_sn.assign(nnzA, 0);
    S.a_pos.assign(nnzA, 0);
    for (int i = 0; i < n; ++i)
        for (int p = Ap[i]; p < Ap[i + 1]; ++p) {
            int a = iperm[i], b = iperm[Aj[p]];
            int col = std::min(a, b), row = std::max(a, b);
            int s = sn_of[col];
            int f = S.fsize[s];
            const int* rr = &S.rows[S.rowptr[s]];
            int pos = static_cast<int>(std::lower_bound(rr, rr + f, row) - rr);
            if (pos >= f || rr[pos] != row) throw std::runtime_error("chol_analyze: A entry outside structure");
            S.a_sn[p] = s;
            S.a_pos[p] = static_cast<int64_t>(col - c0[s]) * f + pos;
        }

    // 11. schedule: leaf subtrees of small supernodes, then levels
    std::vector<double> work(ns_total), sub(ns_total);
    std::vector<char> small(ns_total), allsmall(ns_total);
    double total = 0;
    for (int s = 0; s < ns_total; ++s) {
        int ns = c0[s + 1] - c0[s];
        work[s] = static_cast<double>(S.fsize[s]) * ns;
        total += work[s];
        small[s] = S.fsize[s] <= opt.small_front && ns <= 128;
    }
    std::vector<int> subn(ns_total, 1);   // supernodes in the subtree
    for (int s = 0; s < ns_total; ++s) { sub[s] = work[s]; allsmall[s] = small[s]; }
    for (int s = 0; s < ns_total; ++s) {   // children precede parents
        int p = S.parent[s];
        if (p >= 0) { sub[p] += sub[s]; subn[p] += subn[s]; if (!allsmall[s]) allsmall[p] = 0; }
    }
    double wleaf = std::max(4096.0, total / opt.leaf_tasks);
    std::vector<char> inleaf(ns_total, 0);
    std::vector<int> level(ns_total, 0);
    S.leaf_ptr.assign(1, 0);
    for (int s = 0; s < ns_total; ++s) {
        bool ok = allsmall[s] && sub[s] <= wleaf && subn[s] <= opt.leaf_max_sn;
        int p = S.parent[s];
        bool pok = p >= 0 && allsmall[p] && sub[p] <= wleaf && subn[p] <= opt.leaf_max_sn;
        if (ok && !pok) {
            // s is a leaf-subtree root: collect its subtree in postorder (contiguous range)
            std::vector<int> st = {s}, out;
            while (!st.empty()) {
                int v = st.back();
                st.pop_back();
                out.push_back(v);
                for (int c : children[v]) st.push_back(c);
            }
            std::sort(out.begin(), out.end());
            for (int v : out) { inleaf[v] = 1; S.leaf_sn.push_back(v); }
            S.leaf_ptr.push_back(static_cast<int>(S.leaf_sn.size()));
        }
    }
    int maxlev = 0;
    for (int s = 0; s < ns_total; ++s) {
        if (inleaf[s]) { level[s] = 0; continue; }
        int l = 1;
        for (int c : children[s]) l = std::max(l, level[c] + 1);
        level[s] = l;
        maxlev = std::max(maxlev, l);
    }
    S.level_ptr.assign(maxlev + 1, 0);
    for (int s = 0; s < ns_total; ++s) if (!inleaf[s]) S.level_ptr[level[s]]++;
    // level_ptr[l-1]..level_ptr[l] holds level l (1-based levels)
    {
        std::vector<int> c(maxlev + 1, 0);
        for (int s = 0; s < ns_total; ++s) if (!inleaf[s]) c[level[s]]++;
        S.level_ptr.assign(maxlev + 1, 0);
        for (int l = 1; l <= maxlev; ++l) S.level_ptr[l] = S.level_ptr[l - 1] + c[l];
        S.level_sn.assign(S.level_ptr[maxlev], 0);
        std::vector<int> fill(S.level_ptr.begin(), S.level_ptr.end());
        for (int s = 0; s < ns_total; ++s)
            if (!inleaf[s]) S.level_sn[fill[level[s] - 1]++] = s;
    }
    S.perm = perm;
    S.iperm = iperm;
}

}  // namespace wrapd
