// Execution plan for the numeric supernodal factorization on the GPU (the numeric
// half of the reference's LinearSolver::factorize, deps/mcloptlib/include/MCL/
// LinearSolver.hpp:279-298, redone whenever reweighting changes W, src/System.cpp:50-63).
//
// The elimination tree is processed level by level (a level holds supernodes whose
// children all sit at lower levels).  Per level:
//   1. assemble: one wave per chunk of rows (the plan's asm_chunk, kAsmChunk by default) of a
//      front column builds it in LDS from A's entries and the children's Schur-complement
//      columns (children in index order, so every entry is summed in a fixed order, no
//      atomics) and stores it once;
//   2. per 64-column panel k of the supernode columns, batched launches:
//        POTRF         factor the panel's diagonal tile L_kk (and invert it); a launch of
//                      its own for panel 0 only: the UPDATE task of panel k - 1 on the tile
//                      (k, k) goes on to factor it
//        TRSM + GDIAG  L[t, k] = F[t, k] L_kk^{-T} for row tiles t > k, and
//                      X[k, c] = L_kk^{-1} X[k, c] for column tiles c < k
//        SYRK + GUPD   F[a, b] -= L[a, k] L[b, k]^T for tiles a >= b, k < b < P (the
//                      supernode's remaining columns only), and X[i, c] -= L[i, k] X[k, c]
//                      for L11 row tiles k < i < P, c <= k
//      after which X = L11^{-1} (the top of G_s).  On levels with many update rows the panel
//      steps cover L11's rows only, and one L21 launch then forms each 64-row tile t of the
//      update rows left-looking, in place: for k < P, F[t, k] -= L[t, p] L[k, p]^T for
//      p = 0, 1, .., k - 1 (each product from zero, subtracted in that order), then
//      L[t, k] = F[t, k] L_kk^{-T}.  These are the right-looking steps' operations in the
//      right-looking order, so L21 comes out bit for bit the same, with the update rows read
//      and written once per row tile instead of once per panel step.
//   3. FINAL launches for the rest, each 128x128 super-tile a single pass over all ns
//      columns, one per panel count: the supernodes whose panels end early are finished on a
//      second stream while the level's wider supernodes go on with their panels:
//        SCHUR         F[a, b] -= sum_q L[a, q] L[b, q]^T for update super-tiles a >= b
//        GBOT          G[a, c] = -sum_q L[a, q] X[q, c] for update rows a, L11 columns c
//      so the Schur complement is read and written once (not once per panel).
//      Every task of a launch writes its own tile and reads only tiles that no task of
//      the same launch writes.  G_s = [L11^{-1}; -L21 L11^{-1}] (f_s x ns_s, column-major).
// Front tiles: the ns columns are cut into P = ceil(ns/64) tiles from 0, the m = f - ns
// update rows into Q = ceil(m/64) tiles from ns; tile t < P starts at 64 t, tile t >= P
// at ns + 64 (t - P).  Fronts are f x f column-major doubles in one pool; a front's
// memory is handed to later levels once its parent has assembled it.
#pragma once

#include <cstdint>
#include <vector>

#include "chol_symbolic.hpp"

namespace wrapd {

// FK_POTRF16: FK_POTRF of supernodes with at most kPotrf16Cols columns (one 16x16 block)
enum FactorKind { FK_ASSEMBLE = 0, FK_POTRF = 1, FK_TRSM = 2, FK_UPDATE = 3, FK_FINAL = 4, FK_L21 = 5, FK_POTRF16 = 6 };
constexpr int kPotrf16Cols = 16;

// task pair (s, y): FK_ASSEMBLE y = j | (q << 16), rows [q C, (q + 1) C) (C = asm_chunk) of
// front column j (from row j on); FK_POTRF, FK_POTRF16 y = 0;
// FK_TRSM y = (kind << 30) | t with kind 0 TRSM (row tile t), 1 GDIAG (column tile t);
// FK_UPDATE / FK_FINAL y = (kind << 30) | (a << 15) | b with kind 0 SYRK / SCHUR
// (tiles a, b), 1 GUPD / GBOT (row tile a, column tile b); FK_L21 y = row tile t >= P.
constexpr int kTaskKindShift = 30;
constexpr int kTaskTileShift = 15;
constexpr int kTaskTileMask = (1 << 15) - 1;
// the kind field of a task word (read unsigned: a kind of 2 or more sets the sign bit)
constexpr int task_kind(int y) { return static_cast<int>(static_cast<unsigned>(y) >> kTaskKindShift); }

struct FactorLaunch {
    int kind;
    int panel;          // k (unused for FK_ASSEMBLE)
    int64_t offset;     // first task (int pairs) in tasks
    int count;          // tasks
    int rows = 0;       // FK_ASSEMBLE: the most rows a task of the launch builds (its LDS column)
    // FK_L21 / FK_FINAL of supernodes whose panels end before the level's last panel step:
    // run on a second stream once the main stream has issued the launches before them, so
    // that they overlap the remaining panel steps of the level's wider supernodes
    bool side = false;
    // FK_FINAL: its place among the level's FINAL launches.  FK_ASSEMBLE: the second
    // stream's work it waits for: -1 all of it; k >= 0 that stream's launches up to and
    // including the previous level's FINAL launch k (the last its fronts' children are in),
    // so that it overlaps the later ones
    int ord = -1;
    int wait = -1;
};

struct FactorPlan {
    int n_levels = 0;
    std::vector<int64_t> foff;      // [nsuper] front offset in the pool (doubles)
    int64_t pool_size = 0;          // doubles, peak of the reuse schedule
    int64_t pool_no_reuse = 0;      // sum f_s^2, for the record
    std::vector<int64_t> cbase;     // [nsuper] first global front column of s
    std::vector<int64_t> acol_ptr;  // [total columns + 1]
    std::vector<int> acol;          // pairs (row in front, A slot)
    std::vector<int64_t> ccol_ptr;  // [total columns + 1]
    std::vector<int> ccol;          // pairs (child supernode, child update column jj)
    int asm_chunk = 0;              // rows per assembly task
    // per child c, for each row chunk q of its parent's front (q = 0 .. nq, nq = the chunk
    // count): the first update row ii of c with rel[ii] >= asm_chunk q (entry nq is m_c)
    std::vector<int64_t> cchunk_ptr;   // [nsuper + 1]
    std::vector<int> cchunk;
    std::vector<int> tasks;         // pairs, all launches concatenated
    std::vector<FactorLaunch> launches;   // in stream order
    double flops = 0;               // dense flops issued by the tile tasks (incl. G)
};

constexpr int kFactorTile = 64;
// rows of a front column per assembly task (C5 refactorization: 1024 rows 63.7 ms, 512 rows
// 62.1 ms, 256 rows 64.0 ms)
constexpr int kAsmChunk = 512;
// FK_FINAL tasks work on T x T output tiles, T = kFinalTile: SCHUR (u, v), u >= v: rows
// ns + T u.., columns ns + T v.. of the update block; GBOT (u, c): rows ns + T u.. of G's
// bottom, columns T c.. (< ns).  (128x128 tiles, with twice the operand reuse, measured
// slower on the C5 refactorization at every workgroup shape tried: 61.2-62.1 vs 58.7 ms.)
constexpr int kFinalTile = 64;
constexpr int kFinalChunk = 16;   // the FINAL kernel's k-chunk
// a level whose right-looking panel steps would add at most this many update-row tiles per
// step on average forms L21 in the panel steps (it is bound by the diagonal tiles' chain)
constexpr double kL21RightMaxTiles = 256;
inline int final_row_tiles(int ns, int f) { return (f - ns + kFinalTile - 1) / kFinalTile; }
inline int final_col_tiles(int ns) { return (ns + kFinalTile - 1) / kFinalTile; }

inline int tile_count_l11(int ns) { return (ns + kFactorTile - 1) / kFactorTile; }
inline int tile_count(int ns, int f) { return tile_count_l11(ns) + (f - ns + kFactorTile - 1) / kFactorTile; }

// The plan's choices (the defaults are the measured best; tests set others to exercise
// every path, which give the same factor bit for bit)
struct PlanOptions {
    double l21_right_max_tiles = kL21RightMaxTiles;   // 0: L21 launches everywhere; huge: never
    int asm_chunk = kAsmChunk;                         // rows per assembly task, 64 .. 1024
    bool asm_groups = true;    // false: one assembly launch per level, after all of the previous level's work
    bool potrf16 = true;       // false: every panel-0 diagonal tile by the 64x64 tile kernel
};

void chol_plan(const CholSymbolic& S, FactorPlan& P, const PlanOptions& opt = PlanOptions());

}  // namespace wrapd
