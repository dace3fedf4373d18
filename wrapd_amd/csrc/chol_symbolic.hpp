// Host symbolic analysis for the supernodal Cholesky of A = sum_e w_e^2 D_e^T D_e
// over the free vertices (the matrix the reference hands to SimplicialLDLT with AMD,
// deps/mcloptlib/include/MCL/LinearSolver.hpp:279-298, assembled by
// src/TetElements.cpp:483-510).
//
// Output layout, shared with the device kernels:
//   * the permutation perm (new -> old) from geometric nested dissection, postordered;
//   * supernodes s with columns [col0[s], col0[s+1]) and a front of f_s rows: its own
//     ns_s columns followed by the m_s rows of its off-diagonal structure (sorted);
//   * per supernode a dense column-major block G_s of f_s x ns_s doubles holding
//     (filled by the numeric factorization, wd_factor.hip, scheduled by chol_plan)
//       G_s = [ L11^{-1} ; -L21 L11^{-1} ]
//     so the forward solve is   [y_J ; u] = G_s c_s   and the backward solve is
//     x_J = G_s^T [y_J ; x_R]  (both single dense products, no in-block recursion);
//   * forward assembly lists: for every front row, the child update entries that
//     add into it (deterministic gather, no atomics);
//   * a schedule: small supernodes grouped into leaf subtrees (one workgroup walks a
//     whole subtree), the remaining supernodes by level.
#pragma once

#include <cstdint>
#include <vector>

namespace wrapd {

struct CholSymbolic {
    int n = 0;
    std::vector<int> perm;        // new -> old
    std::vector<int> iperm;       // old -> new
    int nsuper = 0;
    std::vector<int> col0;        // [nsuper+1]
    std::vector<int> fsize;       // front rows f_s
    std::vector<int64_t> rowptr;  // [nsuper+1] into rows
    std::vector<int> rows;        // front rows (new indices), first ns are the columns
    std::vector<int> parent;      // supernodal etree
    std::vector<int64_t> goff;    // [nsuper+1] offsets of G_s (f_s * ns_s doubles each)
    std::vector<int64_t> uoff;    // [nsuper+1] offsets of u_s in rows (m_s rows each)
    // forward assembly: for global front row index p (rowptr[s] + r), the u-rows adding into it
    std::vector<int64_t> asm_ptr; // [rows.size()+1]
    std::vector<int64_t> asm_idx; // u-row indices
    // schedule
    std::vector<int> leaf_ptr, leaf_sn;    // leaf subtrees, supernodes in postorder
    std::vector<int> level_ptr, level_sn;  // top supernodes grouped by level (bottom-up)
    int max_front = 0, max_ns = 0;
    double nnz_L = 0, flops_factor = 0, g_entries = 0;
    double flops_g = 0;   // forming G's blocks: L11^{-1} (ns^3/3) and -L21 L11^{-1} (m ns^2) per supernode

    // A (upper triangle CSR over old free indices) -> lower entries in new numbering:
    // for each A slot the supernode column position, used to scatter A into fronts.
    std::vector<int> a_sn;        // supernode owning the slot's column
    std::vector<int64_t> a_pos;   // column-major offset inside that supernode's front (f_s x f_s)
    // update extend-add maps: child c's front rows ns_c.. map to parent positions
    std::vector<int64_t> relptr;  // [nsuper+1] into rel (m_s entries each)
    std::vector<int> rel;         // position of each off-diagonal row of s inside parent's front
};

// G_s is stored column-major with its front size f rounded up to 16 rows (128 B): tasks
// that read 16-row segments of a column then touch whole cache lines.
constexpr int kGAlign = 16;
constexpr int g_ld(int f) { return (f + kGAlign - 1) / kGAlign * kGAlign; }

struct SymbolicOptions {
    int leaf_size = 48;          // nested-dissection leaf size
    int ns_max = 1024;           // widest supernode after splitting
    int relax_small = 16;        // always merge a child into its parent below this width
    double relax_frac = 0.06;    // allowed fraction of explicit zeros for larger merges
    int relax_max = 1024;        // no relaxed merge beyond this width (wide merges near the root shorten the solves' dependency chains)
    int small_front = 256;       // leaf-subtree supernodes must have f_s <= this
    int leaf_tasks = 2048;       // target number of leaf subtrees
    int leaf_max_sn = 6;         // most supernodes in one leaf subtree (walked one after another by one workgroup)
};

// Ap/Aj: upper-triangular CSR (row i holds columns j >= i) over n free vertices;
// xyz: (n,3) coordinates used for the dissection, or null for a dissection by
// breadth-first level structures of the matrix graph.
void chol_analyze(int n, const std::vector<int>& Ap, const std::vector<int>& Aj, const double* xyz,
                  const SymbolicOptions& opt, CholSymbolic& out);

}  // namespace wrapd
