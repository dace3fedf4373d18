// C ABI (include/wrapd_gpu.h): context management, mesh upload, and the host control of
// the device work.  The control flow of the global step follows SplitLBFGS::minimize
// (deps/mcloptlib/include/MCL/SplitLBFGS.hpp:114-256) decision for decision; only the
// vector arithmetic runs on the device.
#include <algorithm>
#include <chrono>
#include <mutex>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "wd_ctx.hpp"

using wd::check_hip;

wrapd_ctx::~wrapd_ctx() {
    if (fev_fork) (void)hipEventDestroy(fev_fork);
    if (fev_join) (void)hipEventDestroy(fev_join);
    for (hipEvent_t e : fev_final) (void)hipEventDestroy(e);
    if (fstream2) (void)hipStreamDestroy(fstream2);
    auto f = [](void* p) { if (p) (void)hipFree(p); };
    f(vidx); f(Di); f(vol); f(kst); f(mu_e); f(lam_e); f(wsq); f(cand); f(U); f(P); f(symF); f(sig); f(gstage);
    f(xall); f(grad); f(grad_old); f(qv); f(rv); f(bv); f(xref);
    for (int i = 0; i < wd::kMaxHist; ++i) { f(hS[i]); f(hY[i]); }
    f(v2e_ptr); f(cpos); f(a_cptr); f(a_cidx); f(Ax); f(fb_info); f(fb_list); f(fb_count); f(fb_tmp); f(d_fin_ticket);
    f(G); f(ysol); f(xsol); f(ubuf); f(d_perm); f(d_iperm); f(d_col0); f(d_fsize); f(d_rowptr); f(d_rows);
    f(d_goff); f(d_uoff); f(d_asm_ptr); f(d_asm_idx); f(d_udest); f(d_leaf_ptr); f(d_leaf_sn); f(d_flow_f); f(d_flow_b);
    f(d_flow_abort); f(d_bcopy); f(bperm);
    f(pool); f(d_foff); f(d_cbase); f(d_acol_ptr); f(d_acol); f(d_ccol_ptr); f(d_ccol); f(d_relptr); f(d_rel);
    f(d_cchunk_ptr); f(d_cchunk);
    f(d_ftasks); f(d_ffail);
    f(mlp_master); f(mlp_m1); f(mlp_m2); f(mlp_gsum); f(mlp_norm); f(mlp_bf); f(mlp_loss); f(mlp_partial);
    f(mlp_lpart); f(mlp_log);
    f(partials); f(scal); f(d_red_abort);
    if (scal_host) (void)hipHostFree(scal_host);
    if (host_seq) (void)hipHostFree(host_seq);
    for (auto& p : prof.pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
    for (auto e : prof.pool) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
}

namespace wd {

int check_hip(wrapd_ctx* c, hipError_t e, const char* what) {
    if (e == hipSuccess) return 0;
    if (c) c->last_error = std::string(what) + ": " + hipGetErrorString(e);
    return 1;
}

static hipEvent_t ev_get(wrapd_ctx* c) {
    if (!c->prof.pool.empty()) {
        hipEvent_t e = c->prof.pool.back();
        c->prof.pool.pop_back();
        return e;
    }
    hipEvent_t e;
    (void)hipEventCreate(&e);
    return e;
}

ProfScope::ProfScope(wrapd_ctx* ctx, const char* name) : c(ctx) {
    if (!c->prof.on || (!c->prof.only.empty() && c->prof.only != name)) return;
    Prof::Pending p;
    p.name = name;
    p.a = ev_get(c);
    p.b = ev_get(c);
    (void)hipEventRecord(p.a, c->stream);
    c->prof.pending.push_back(p);
    idx = static_cast<int>(c->prof.pending.size()) - 1;
}

ProfScope::~ProfScope() {
    if (idx >= 0) (void)hipEventRecord(c->prof.pending[idx].b, c->stream);
}

static void prof_flush(wrapd_ctx* c) {
    if (c->prof.pending.empty()) return;
    (void)hipStreamSynchronize(c->stream);
    for (auto& p : c->prof.pending) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, p.a, p.b);
        auto& a = c->prof.acc[p.name];
        a.first += ms;
        a.second += 1.0;
        c->prof.pool.push_back(p.a);
        c->prof.pool.push_back(p.b);
    }
    c->prof.pending.clear();
}

template <typename T>
static int dalloc(wrapd_ctx* c, T*& p, size_t n) {
    if (p) { (void)hipFree(p); p = nullptr; }
    return check_hip(c, hipMalloc(&p, std::max<size_t>(1, n) * sizeof(T)), "hipMalloc");
}

template <typename T>
static int h2d(wrapd_ctx* c, T* d, const T* h, size_t n) {
    if (n == 0) return 0;
    return check_hip(c, hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, c->stream), "H2D");
}

template <typename T>
static int d2h(wrapd_ctx* c, T* h, const T* d, size_t n) {
    if (n == 0) return 0;
    if (check_hip(c, hipMemcpyAsync(h, d, n * sizeof(T), hipMemcpyDeviceToHost, c->stream), "D2H")) return 1;
    return check_hip(c, hipStreamSynchronize(c->stream), "sync");
}

// read n device scalars starting at slot `at`.  Every slot the host reads is written, by
// the reduction that produces it, to the pinned copy too, followed by that reduction's
// epoch in host_seq: the host waits for the latest epoch by polling host memory, and falls
// back to a stream synchronize (which also reports a failed launch) after about 50 ms.
// `want`: wait for that reduction (host_seq is the epoch of the last one stored; epochs only
// grow, so a later one implies it).
static const double* scal_read_epoch(wrapd_ctx* c, int at, unsigned want) {
    auto arrived = [&] { return static_cast<int>(__atomic_load_n(c->host_seq, __ATOMIC_ACQUIRE) - want) >= 0; };
    const auto t0 = std::chrono::steady_clock::now();
    for (unsigned spins = 0;; ++spins) {
        if (arrived()) return c->scal_host + at;
        if ((spins & 1023) == 1023 &&
            std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(50))
            break;
    }
    if (check_hip(c, hipStreamSynchronize(c->stream), "sync")) return nullptr;
    if (!arrived()) {
        c->last_error = "device scalars: the reduction's host copy did not arrive";
        return nullptr;
    }
    return c->scal_host + at;
}
// A reduction whose wait for a block's partials ran out leaves NaN results and sets the
// sticky abort word; the partials are then in no defined state, so the context refuses all
// further work (wrapd_destroy still works).
static bool reduction_aborted(wrapd_ctx* c) {
    int v = 0;
    if (hipMemcpyAsync(&v, c->d_red_abort, sizeof(int), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess || v != 0) {
        c->dead = true;
        c->last_error = "reduction: a block's partial results did not arrive (wait timed out); the context is unusable";
        return true;
    }
    return false;
}

static const double* scal_read(wrapd_ctx* c, int at, int n) {
    const double* r = scal_read_epoch(c, at, c->host_epoch);
    if (!r) return r;
    for (int k = 0; k < n; ++k)
        if (std::isnan(r[k]) && reduction_aborted(c)) return nullptr;
    return r;
}

// Entry points that change what a begun global step's first launches read (x, the duals,
// the weights, the history, the factorization) drop it: the global step then launches
// them again itself.
static void drop_begun(wrapd_ctx* c) { c->gs_begun = false; }

static int launch_status(wrapd_ctx* c) {
    return check_hip(c, hipGetLastError(), "kernel launch") ? WRAPD_ERR_HIP : WRAPD_OK;
}


}  // namespace wd

// ---- the static scratch array of the allocation-free entry points ------------------------
// wrapd_mlp_grad_noalloc stages its tuples, sample list, per-wave partial gradients and sums
// here instead of in memory it allocates.  Every GPU holds its own copy of the array until
// the process ends, at its own device address: the host never passes `scratch` itself to a
// kernel or a copy (in host code the name is a host-side stand-in), only the address
// hipGetSymbolAddress returns with that device current, looked up once per device at the
// first wrapd_create on it.  All calls on one GPU share its array, and its contents stay from
// call to call: calls that overlap on one GPU (several contexts or host threads) race.
namespace {
constexpr size_t kScratchBytes = size_t(8) << 20;   // 8 MiB per GPU
__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;   // per device index, null until looked up

int scratch_lookup(wrapd_ctx* c) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (static_cast<int>(g_scratch_ptr.size()) <= c->device) g_scratch_ptr.resize(c->device + 1, nullptr);
    if (!g_scratch_ptr[c->device]) {
        void* p = nullptr;   // c->device is current (wrapd_create set it)
        if (wd::check_hip(c, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)") || !p)
            return WRAPD_ERR_HIP;
        g_scratch_ptr[c->device] = p;
    }
    c->scratch = g_scratch_ptr[c->device];
    return WRAPD_OK;
}

// wrapd_mlp_grad_noalloc's layout inside the scratch array, for n samples
struct MlpGradScratch {
    size_t partial, lpart, gsum, loss, tuples, idx, end;
};
constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }
constexpr MlpGradScratch mlp_grad_layout(long long n) {
    const size_t waves = static_cast<size_t>((n + 255) / 256);
    MlpGradScratch L{};
    L.partial = 0;
    L.lpart = align256(L.partial + waves * wd::kMlpParams * sizeof(float));
    L.gsum = align256(L.lpart + waves * sizeof(float));
    L.loss = align256(L.gsum + wd::kMlpParams * sizeof(float));
    L.tuples = align256(L.loss + sizeof(float));
    L.idx = align256(L.tuples + 9ull * n * sizeof(float));
    L.end = align256(L.idx + static_cast<size_t>(n) * sizeof(int));
    return L;
}
constexpr long long kMlpGradNoAllocMax = 65536;
static_assert(mlp_grad_layout(kMlpGradNoAllocMax).end <= kScratchBytes, "scratch too small for the stated maximum");
}  // namespace

using namespace wd;

#define CTX_OK(c) do {                                                                        \
        if (!(c)) return WRAPD_ERR_ARG;                                                       \
        if ((c)->dead) {                                                                      \
            (c)->last_error = "the context is unusable after a device wait timed out; destroy it"; \
            return WRAPD_ERR_STATE;                                                           \
        }                                                                                     \
    } while (0)
#define MESH_OK(c) do { if (!(c)->mesh_ready) { (c)->last_error = "mesh not uploaded"; return WRAPD_ERR_STATE; } } while (0)

extern "C" {

int wrapd_create(wrapd_ctx** out, int device) {
    if (!out) return WRAPD_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return WRAPD_ERR_NO_DEVICE;
    wrapd_ctx* c = new wrapd_ctx();
    c->device = device;
    // The host waits for a few scalars several times per ADMM iteration: spin rather than
    // yield, which shortens each wait's wake-up.  Only possible before the device's
    // context exists (otherwise the process keeps its setting).
    (void)hipSetDeviceFlags(hipDeviceScheduleSpin);
    (void)hipGetLastError();
    if (check_hip(c, hipSetDevice(device), "hipSetDevice") ||
        check_hip(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate") ||
        dalloc(c, c->partials, static_cast<size_t>(kMaxBlocks) * kRedStride) || dalloc(c, c->scal, S_NUM) ||
        dalloc(c, c->d_red_abort, 1) || dalloc(c, c->d_fin_ticket, 1) ||
        check_hip(c, hipHostMalloc(reinterpret_cast<void**>(&c->scal_host), S_NUM * sizeof(double)), "hipHostMalloc") ||
        check_hip(c, hipHostMalloc(reinterpret_cast<void**>(&c->host_seq), 64), "hipHostMalloc")) {
        delete c;
        return WRAPD_ERR_HIP;
    }
    // partials start pending (each 32-bit half of the pattern is 0xFFF40000)
    static_assert(wd::kPendingBits == 0xFFF40000FFF40000ull, "pending pattern: two equal halves");
    if (check_hip(c, hipMemsetD32(reinterpret_cast<hipDeviceptr_t>(c->partials), 0xFFF40000u,
                                  2 * static_cast<size_t>(kMaxBlocks) * kRedStride), "hipMemsetD32") ||
        check_hip(c, hipMemset(c->d_red_abort, 0, sizeof(int)), "hipMemset") ||
        check_hip(c, hipMemset(c->d_fin_ticket, 0, sizeof(unsigned)), "hipMemset")) {
        delete c;
        return WRAPD_ERR_HIP;
    }
    if (scratch_lookup(c)) {
        delete c;
        return WRAPD_ERR_HIP;
    }
    c->tail_max_blocks = kTailMaxBlocks;
    if (const char* e = std::getenv("WRAPD_TAIL_MAX_BLOCKS")) {
        const int v = std::atoi(e);
        if (v >= 1 && v <= kTailMaxBlocks) c->tail_max_blocks = v;
    }
    if (const char* e = std::getenv("WRAPD_ASM_CHUNK")) {
        const int v = std::atoi(e);
        if (v >= 64 && v <= 1024) c->plan_opt.asm_chunk = v;
    }

    if (const char* e = std::getenv("WRAPD_L21_RIGHT_MAX")) {
        const double v = std::atof(e);
        if (v >= 0.0) c->plan_opt.l21_right_max_tiles = v;
    }
    if (const char* e = std::getenv("WRAPD_ASM_GROUPS")) c->plan_opt.asm_groups = std::atoi(e) != 0;
    if (const char* e = std::getenv("WRAPD_POTRF16")) c->plan_opt.potrf16 = std::atoi(e) != 0;
    if (const char* e = std::getenv("WRAPD_LEARNED_GATE")) {
        const double v = std::atof(e);
        if (v > 0.0) c->learned_gate = v;
    }
    launch_fill(c, c->scal, 0.0, S_NUM);
    std::fill(c->scal_host, c->scal_host + S_NUM, 0.0);
    *c->host_seq = 0;
    *out = c;
    return WRAPD_OK;
}

int wrapd_destroy(wrapd_ctx* c) {
    if (!c) return WRAPD_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    delete c;
    return WRAPD_OK;
}

const char* wrapd_last_error(wrapd_ctx* c) { return c ? c->last_error.c_str() : "null context"; }

int wrapd_sizes(wrapd_ctx* c, int* nv, int* ne, int* nf, int* nfix) {
    CTX_OK(c);
    if (nv) *nv = c->n_v;
    if (ne) *ne = c->n_e;
    if (nf) *nf = c->n_f;
    if (nfix) *nfix = c->n_fix;
    return WRAPD_OK;
}

int wrapd_upload_mesh(wrapd_ctx* c, int n_v, const double* rest, int n_e, const int* tets, int n_pins,
                      const int* pins, const wrapd_material* mat, const double* mu_e, const double* lambda_e) {
    CTX_OK(c);
    drop_begun(c);
    if (n_v <= 0 || n_e <= 0 || !rest || !tets || !mat || n_pins < 0 || (n_pins > 0 && !pins)) {
        c->last_error = "wrapd_upload_mesh: bad arguments";
        return WRAPD_ERR_ARG;
    }
    if (mat->model != 0 && mat->model != 1) { c->last_error = "unknown model"; return WRAPD_ERR_ARG; }
    for (long long i = 0; i < 4ll * n_e; ++i)
        if (tets[i] < 0 || tets[i] >= n_v) { c->last_error = "tet index out of range"; return WRAPD_ERR_ARG; }
    if (blocks_for(n_e, kLocalThreads) > kMaxBlocks || blocks_for(3ll * n_v, kElemThreads) > kMaxBlocks) {
        c->last_error = "wrapd_upload_mesh: mesh larger than the reduction buffers (65536 blocks)";
        return WRAPD_ERR_ARG;
    }
    (void)hipSetDevice(c->device);
    c->n_v = n_v;
    c->n_e = n_e;
    c->n_ep = ((n_e + 255) / 256) * 256;
    c->mat = *mat;
    c->tets.assign(tets, tets + 4ll * n_e);
    c->rest.assign(rest, rest + 3ll * n_v);
    // binding::add_tetmesh (samples/utils/AddMeshes.hpp:89-110)
    c->pins.assign(pins, pins + n_pins);
    std::sort(c->pins.begin(), c->pins.end());
    for (size_t i = 0; i < c->pins.size(); ++i)
        if (c->pins[i] < 0 || c->pins[i] >= n_v || (i > 0 && c->pins[i] == c->pins[i - 1])) {
            c->last_error = "pins out of range or repeated";
            return WRAPD_ERR_ARG;
        }
    c->free_idx.assign(n_v, -1);
    c->fix_idx.assign(n_v, -1);
    int j = 0, k = 0;
    for (int i = 0; i < n_v; ++i) {
        if (j < n_pins && i == c->pins[j]) { c->fix_idx[i] = j; j++; }
        else { c->free_idx[i] = k; k++; }
    }
    c->n_f = k;
    c->n_fix = j;
    if (c->n_f == 0) { c->last_error = "no free vertices"; return WRAPD_ERR_ARG; }
    const int n_ep = c->n_ep, n_f = c->n_f;
    // the tets (padding rows: vertex 0) and each vertex's combined position index go to the
    // device, which builds vidx, A's pattern and contribution lists and the vertex -> corner
    // lists (build_mesh_maps)
    int* d_tv = nullptr;
    int* d_comb = nullptr;
    double* d_rest = nullptr;
    int* d_bad = nullptr;
    auto free_tmp = [&] {
        (void)hipFree(d_tv);
        (void)hipFree(d_comb);
        (void)hipFree(d_rest);
        (void)hipFree(d_bad);
    };
    std::vector<int> comb(n_v);
    for (int i = 0; i < n_v; ++i) comb[i] = c->free_idx[i] >= 0 ? c->free_idx[i] : n_f + c->fix_idx[i];
    int rc = 0;
    rc |= dalloc(c, d_tv, 4ull * n_ep);
    rc |= dalloc(c, d_comb, n_v);
    rc |= dalloc(c, d_rest, 3ull * n_v);
    rc |= dalloc(c, d_bad, 1);
    if (rc) { free_tmp(); return WRAPD_ERR_HIP; }
    rc |= check_hip(c, hipMemsetAsync(d_tv + 4ll * n_e, 0, 4ull * (n_ep - n_e) * sizeof(int), c->stream), "memset");
    rc |= h2d(c, d_tv, tets, 4ull * n_e);
    rc |= h2d(c, d_comb, comb.data(), comb.size());
    rc |= h2d(c, d_rest, rest, 3ull * n_v);
    long long n_corners = 0;
    if (rc || build_mesh_maps(c, d_tv, d_comb, &n_corners) != WRAPD_OK) { free_tmp(); return WRAPD_ERR_HIP; }
    // device allocations
    rc |= dalloc(c, c->Di, 12ull * n_ep);
    rc |= dalloc(c, c->vol, n_ep);
    rc |= dalloc(c, c->kst, n_ep);
    rc |= dalloc(c, c->wsq, n_ep);
    rc |= dalloc(c, c->cand, n_ep);
    rc |= dalloc(c, c->U, 9ull * n_ep);
    rc |= dalloc(c, c->P, 9ull * n_ep);
    rc |= dalloc(c, c->symF, 9ull * n_ep);
    rc |= dalloc(c, c->sig, 3ull * n_ep);
    rc |= dalloc(c, c->gstage, 3ull * n_corners);
    rc |= dalloc(c, c->xall, 3ull * (n_f + c->n_fix));
    rc |= dalloc(c, c->grad, 3ull * n_f);
    rc |= dalloc(c, c->grad_old, 3ull * n_f);
    rc |= dalloc(c, c->qv, 3ull * n_f);
    rc |= dalloc(c, c->rv, 3ull * n_f);
    rc |= dalloc(c, c->bv, 3ull * n_f);
    for (int i = 0; i < kMaxHist; ++i) { rc |= dalloc(c, c->hS[i], 3ull * n_f); rc |= dalloc(c, c->hY[i], 3ull * n_f); }
    rc |= dalloc(c, c->Ax, c->nnzA);
    rc |= dalloc(c, c->fb_info, n_ep);
    rc |= dalloc(c, c->fb_list, n_ep);
    rc |= dalloc(c, c->fb_count, 1);
    if (mu_e) { rc |= dalloc(c, c->mu_e, n_ep); } else if (c->mu_e) { (void)hipFree(c->mu_e); c->mu_e = nullptr; }
    if (lambda_e) { rc |= dalloc(c, c->lam_e, n_ep); } else if (c->lam_e) { (void)hipFree(c->lam_e); c->lam_e = nullptr; }
    if (rc || prepare_local_fallback(c) != WRAPD_OK) { free_tmp(); return WRAPD_ERR_HIP; }
    std::vector<double> me, le;
    if (mu_e) { me.assign(n_ep, mat->mu); std::copy(mu_e, mu_e + n_e, me.begin()); rc |= h2d(c, c->mu_e, me.data(), n_ep); }
    if (lambda_e) { le.assign(n_ep, mat->lambda); std::copy(lambda_e, lambda_e + n_e, le.begin()); rc |= h2d(c, c->lam_e, le.data(), n_ep); }
    c->hetero = mu_e || lambda_e;
    // setup kernel on the rest positions
    int zero = 0;
    rc |= h2d(c, d_bad, &zero, 1);
    if (rc) { free_tmp(); return WRAPD_ERR_HIP; }
    launch_setup(c, d_tv, d_rest, d_bad);
    int bad = 0;
    rc |= launch_status(c);
    rc |= d2h(c, &bad, d_bad, 1);
    free_tmp();
    if (rc) return WRAPD_ERR_HIP;
    // zero the dual state so that reads before init_duals are defined
    launch_fill(c, c->U, 0.0, 9ll * n_ep);
    launch_fill(c, c->P, 0.0, 9ll * n_ep);
    launch_fill(c, c->symF, 0.0, 9ll * n_ep);
    launch_fill(c, c->cand, 1.0, n_ep);
    launch_fill(c, c->grad_old, 0.0, 3ll * n_f);
    c->analyzed = c->factored = false;
    c->m_k = 0;
    for (int i = 0; i < kMaxHist; ++i) c->pos2slot[i] = i;
    if (bad) { c->last_error = "Inverted tet in the rest mesh"; return WRAPD_ERR_INVERTED_REST; }
    c->mesh_ready = true;
    return launch_status(c);
}

int wrapd_set_x_fix(wrapd_ctx* c, const double* x_fix) {
    CTX_OK(c); MESH_OK(c);
    drop_begun(c);
    if (c->n_fix == 0) return WRAPD_OK;
    if (!x_fix) return WRAPD_ERR_ARG;
    return h2d(c, c->xall + 3ll * c->n_f, x_fix, 3ull * c->n_fix) ? WRAPD_ERR_HIP : WRAPD_OK;
}

int wrapd_set_x_free(wrapd_ctx* c, const double* x) {
    CTX_OK(c); MESH_OK(c);
    drop_begun(c);
    if (!x) return WRAPD_ERR_ARG;
    return h2d(c, c->xall, x, 3ull * c->n_f) ? WRAPD_ERR_HIP : WRAPD_OK;
}

int wrapd_get_x_free(wrapd_ctx* c, double* x) {
    CTX_OK(c); MESH_OK(c);
    if (!x) return WRAPD_ERR_ARG;
    return d2h(c, x, c->xall, 3ull * c->n_f) ? WRAPD_ERR_HIP : WRAPD_OK;
}

int wrapd_set_x_ref(wrapd_ctx* c, const double* x) {
    CTX_OK(c); MESH_OK(c);
    if (!x) return WRAPD_ERR_ARG;
    if (!c->xref && dalloc(c, c->xref, 3ull * c->n_f)) return WRAPD_ERR_HIP;
    return h2d(c, c->xref, x, 3ull * c->n_f) ? WRAPD_ERR_HIP : WRAPD_OK;
}

int wrapd_x_distance(wrapd_ctx* c, double* dist) {
    int st = wrapd_x_distance_launch(c);
    if (st != WRAPD_OK) return st;
    return wrapd_x_distance_result(c, dist);
}

int wrapd_x_distance_launch(wrapd_ctx* c) {
    CTX_OK(c); MESH_OK(c);
    if (!c->xref) { c->last_error = "no reference positions"; return WRAPD_ERR_STATE; }
    launch_diffnorm(c, c->xall, c->xref, S_XD);
    c->xd_epoch = c->host_epoch;
    return launch_status(c);
}

int wrapd_x_distance_result(wrapd_ctx* c, double* dist) {
    CTX_OK(c); MESH_OK(c);
    const double* r = scal_read_epoch(c, S_XD, c->xd_epoch);
    if (!r || (std::isnan(r[0]) && reduction_aborted(c))) return WRAPD_ERR_HIP;
    if (dist) *dist = std::sqrt(r[0]);
    return WRAPD_OK;
}

int wrapd_set_rot_aware(wrapd_ctx* c, int rot_aware) {
    CTX_OK(c);
    drop_begun(c);
    c->rot_aware = rot_aware ? 1 : 0;
    return WRAPD_OK;
}

int wrapd_init_duals(wrapd_ctx* c, int rot_aware, double* polar_pe) {
    CTX_OK(c); MESH_OK(c);
    drop_begun(c);
    c->rot_aware = rot_aware ? 1 : 0;
    arm_tail(c, 1, 0u, S_RED);
    launch_init_duals(c, c->rot_aware);
    const double* r = scal_read(c, S_RED, 1);
    if (!r) return WRAPD_ERR_HIP;
    if (polar_pe) *polar_pe = r[0];
    return launch_status(c);
}

int wrapd_defo_energy(wrapd_ctx* c, double* energy, int* flips) {
    CTX_OK(c); MESH_OK(c);
    drop_begun(c);
    arm_tail(c, 3, 0u, S_RED);
    launch_defo(c, c->rot_aware ? (DM_WRITE_SYMF | DM_ENERGY) : (DM_FT_CACHE | DM_ENERGY), 0);
    const double* r = scal_read(c, S_RED, 3);
    if (!r) return WRAPD_ERR_HIP;
    c->last_energy = r[1];
    c->last_flips = static_cast<int>(r[2]);
    if (energy) *energy = r[1];
    if (flips) *flips = static_cast<int>(r[2]);
    return launch_status(c);
}

int wrapd_candidate_weights(wrapd_ctx* c, int use_hessian, int direct) {
    CTX_OK(c); MESH_OK(c);
    drop_begun(c);
    if (c->mat.model != 1) { c->last_error = "Adaptive weights are only supported for the NH model."; return WRAPD_ERR_STATE; }
    launch_defo(c, direct ? (DM_CAND | DM_DIRECT) : DM_CAND, use_hessian);
    return launch_status(c);
}

int wrapd_max_wsq_ratio(wrapd_ctx* c, double* ratio) {
    CTX_OK(c); MESH_OK(c);
    arm_tail(c, 1, 1u, S_RED);
    launch_max_ratio(c);
    const double* r = scal_read(c, S_RED, 1);
    if (!r) return WRAPD_ERR_HIP;
    if (ratio) *ratio = r[0];
    return launch_status(c);
}

int wrapd_update_actual_weights(wrapd_ctx* c) {
    CTX_OK(c); MESH_OK(c);
    drop_begun(c);
    launch_update_weights(c);
    return launch_status(c);
}

int wrapd_local_step_launch(wrapd_ctx* c, const wrapd_local_args* a) {
    CTX_OK(c); MESH_OK(c);
    drop_begun(c);
    if (!a) return WRAPD_ERR_ARG;
    if (a->update_cand && c->mat.model != 1) {
        c->last_error = "Adaptive weights are only supported for the NH model.";
        return WRAPD_ERR_STATE;
    }
    if (a->prox_mode == WRAPD_PROX_LEARNED && !c->mlp_ready) {
        c->last_error = "learned prox: no network (wrapd_mlp_init / wrapd_mlp_set_params)";
        return WRAPD_ERR_STATE;
    }
    arm_tail(c, 11, (1u << 3) | (1u << 9), S_LOCAL);   // a slot range of its own: read after later launches
    launch_local(c, a);
    c->local_epoch = c->host_epoch;
    c->local_pending = true;
    return launch_status(c);
}

int wrapd_local_step_result(wrapd_ctx* c, double pe_prev, wrapd_local_out* out) {
    CTX_OK(c);
    if (!c->local_pending) {
        c->last_error = "local step results: no local step launched";
        return WRAPD_ERR_STATE;
    }
    c->local_pending = false;
    const double* r0 = scal_read_epoch(c, S_LOCAL, c->local_epoch);
    if (!r0) return WRAPD_ERR_HIP;
    // an FP32 step's sums are those of its element kernel plus those of its FP64 pass over
    // the deferred elements (S_LOCAL_FB, reduced after it: the epoch waited for is that one)
    double r[11];
    for (int k = 0; k < 11; ++k) {
        const double fb = c->local_fb ? r0[S_LOCAL_FB - S_LOCAL + k] : 0.0;
        r[k] = (k == 3 || k == 9) ? std::max(r0[k], fb) : r0[k] + fb;
    }
    if (out) {
        out->bef = r[0];
        out->aft = r[1];
        out->pe = r[2];
        out->max_wsq_ratio = r[3];
        out->prox_iters_total = r[4];
        out->fallback_count = r[5];
        out->neg_sigma_count = r[6];
        out->learned_accepted = r[7];
        out->evals_total = r[8];
        out->evals_max = r[9];
        out->learned_used = r[10];
        out->delta_auglag = r[1] - r[0] + r[2] - pe_prev;
    }
    int st = launch_status(c);
    if (st == WRAPD_OK && r[6] > 0) {
        c->last_error = "Negative singular value in solution";
        return WRAPD_ERR_NEG_SIGMA;
    }
    return st;
}

int wrapd_local_step(wrapd_ctx* c, const wrapd_local_args* a, double pe_prev, wrapd_local_out* out) {
    const int st = wrapd_local_step_launch(c, a);
    return st != WRAPD_OK ? st : wrapd_local_step_result(c, pe_prev, out);
}

int wrapd_local_step_unaware(wrapd_ctx* c, const wrapd_local_args* a) {
    CTX_OK(c); MESH_OK(c);
    drop_begun(c);
    if (!a) return WRAPD_ERR_ARG;
    launch_local_unaware(c, a);
    return launch_status(c);
}

int wrapd_global_value_grad(wrapd_ctx* c, double* value, double* grad) {
    CTX_OK(c); MESH_OK(c);
    drop_begun(c);
    arm_tail(c, 1, 0u, S_VAL);
    launch_defo(c, DM_GRAD | DM_VALUE, 0);
    launch_gather_grad(c, c->grad, 0, 0, nullptr, 0);
    const double* r = scal_read(c, S_VAL, 1);
    if (!r) return WRAPD_ERR_HIP;
    if (value) *value = r[0];
    if (grad && d2h(c, grad, c->grad, 3ull * c->n_f)) return WRAPD_ERR_HIP;
    return launch_status(c);
}

int wrapd_refactor(wrapd_ctx* c) {
    CTX_OK(c); MESH_OK(c);
    drop_begun(c);
    launch_assemble_A(c);
    if (!c->analyzed) {
        std::vector<double> xyz(3ull * c->n_f);
        for (int v = 0; v < c->n_v; ++v) {
            const int f = c->free_idx[v];
            if (f >= 0) for (int q = 0; q < 3; ++q) xyz[3 * f + q] = c->rest[3 * v + q];
        }
        wrapd::SymbolicOptions opt;
        try {
            wrapd::chol_analyze(c->n_f, c->Ap, c->Aj, xyz.data(), opt, c->sym);
        } catch (const std::exception& ex) {
            c->last_error = ex.what();
            return WRAPD_ERR_STATE;
        }
        int st = chol_prepare_device(c);
        if (!st) st = factor_prepare_device(c);
        if (st) return st;
        c->analyzed = true;
    }
    {
        ProfScope ps(c, "refactor");
        launch_factor(c);
    }
    int fail = 0;
    if (check_hip(c, hipMemcpyAsync(&fail, c->d_ffail, sizeof(int), hipMemcpyDeviceToHost, c->stream), "factor status") ||
        check_hip(c, hipStreamSynchronize(c->stream), "factor"))
        return WRAPD_ERR_HIP;
    if (fail) {
        c->last_error = "factorization: non-positive pivot (A not positive definite)";
        c->factored = false;
        return WRAPD_ERR_NOT_SPD;
    }
    c->factored = true;
    return launch_status(c);
}

int wrapd_linsolve_analyze(wrapd_ctx* c, int n, const int* row_ptr, const int* cols, const double* xyz) {
    CTX_OK(c);
    if (c->mesh_ready) { c->last_error = "wrapd_linsolve_analyze: this context holds a mesh"; return WRAPD_ERR_STATE; }
    if (n <= 0 || !row_ptr || !cols || row_ptr[0] != 0) { c->last_error = "wrapd_linsolve_analyze: bad arguments"; return WRAPD_ERR_ARG; }
    for (int i = 0; i < n; ++i) {
        // upper triangle by rows, columns strictly increasing, the diagonal first
        if (row_ptr[i + 1] <= row_ptr[i] || cols[row_ptr[i]] != i) {
            c->last_error = "wrapd_linsolve_analyze: row " + std::to_string(i) + " lacks its diagonal entry";
            return WRAPD_ERR_ARG;
        }
        for (int p = row_ptr[i] + 1; p < row_ptr[i + 1]; ++p)
            if (cols[p] <= cols[p - 1] || cols[p] >= n) {
                c->last_error = "wrapd_linsolve_analyze: row " + std::to_string(i) +
                                " is not sorted upper-triangular with columns < n";
                return WRAPD_ERR_ARG;
            }
    }
    (void)hipSetDevice(c->device);
    drop_begun(c);
    c->n_f = n;
    c->n_v = n;
    c->Ap.assign(row_ptr, row_ptr + n + 1);
    c->Aj.assign(cols, cols + row_ptr[n]);
    c->nnzA = row_ptr[n];
    if (dalloc(c, c->Ax, c->nnzA) || dalloc(c, c->bv, 3ull * n) || dalloc(c, c->rv, 3ull * n)) return WRAPD_ERR_HIP;
    wrapd::SymbolicOptions opt;
    try {
        wrapd::chol_analyze(n, c->Ap, c->Aj, xyz, opt, c->sym);
    } catch (const std::exception& ex) {
        c->last_error = ex.what();
        return WRAPD_ERR_STATE;
    }
    int st = chol_prepare_device(c);
    if (!st) st = factor_prepare_device(c);
    if (st) return st;
    c->linsolve = true;
    c->analyzed = true;
    c->factored = false;
    return launch_status(c);
}

int wrapd_linsolve_factorize(wrapd_ctx* c, const double* vals) {
    CTX_OK(c);
    if (!c->linsolve || !c->analyzed) { c->last_error = "wrapd_linsolve_factorize: no analyzed pattern"; return WRAPD_ERR_STATE; }
    if (!vals) return WRAPD_ERR_ARG;
    if (h2d(c, c->Ax, vals, static_cast<size_t>(c->nnzA))) return WRAPD_ERR_HIP;
    {
        ProfScope ps(c, "refactor");
        launch_factor(c);
    }
    int fail = 0;
    if (check_hip(c, hipMemcpyAsync(&fail, c->d_ffail, sizeof(int), hipMemcpyDeviceToHost, c->stream), "factor status") ||
        check_hip(c, hipStreamSynchronize(c->stream), "factor"))
        return WRAPD_ERR_HIP;
    c->factored = !fail;
    if (fail) {
        c->last_error = "factorization: non-positive pivot (A not positive definite)";
        return WRAPD_ERR_NOT_SPD;
    }
    return launch_status(c);
}

int wrapd_solve3(wrapd_ctx* c, const double* b, double* x) {
    CTX_OK(c);
    if (!c->mesh_ready && !c->linsolve) { c->last_error = "mesh not uploaded"; return WRAPD_ERR_STATE; }
    drop_begun(c);
    if (!c->factored) { c->last_error = "no factorization"; return WRAPD_ERR_STATE; }
    if (!b || !x) return WRAPD_ERR_ARG;
    if (h2d(c, c->bv, b, 3ull * c->n_f)) return WRAPD_ERR_HIP;
    launch_solve(c, c->bv, c->rv);
    if (d2h(c, x, c->rv, 3ull * c->n_f)) return WRAPD_ERR_HIP;
    if (solve_aborted(c)) { c->last_error = "triangular solve: a dependency wait timed out"; return WRAPD_ERR_HIP; }
    return launch_status(c);
}

int wrapd_get_A(wrapd_ctx* c, int* row_ptr, int* cols, double* vals) {
    CTX_OK(c); MESH_OK(c);
    if (row_ptr) std::copy(c->Ap.begin(), c->Ap.end(), row_ptr);
    if (cols) std::copy(c->Aj.begin(), c->Aj.end(), cols);
    if (vals) {
        launch_assemble_A(c);
        if (d2h(c, vals, c->Ax, c->nnzA)) return -WRAPD_ERR_HIP;
    }
    return c->nnzA;
}

int wrapd_chol_stats(wrapd_ctx* c, double* st) {
    CTX_OK(c);
    if (!c->analyzed) return WRAPD_ERR_STATE;
    st[0] = c->sym.nsuper;
    st[1] = c->sym.nnz_L;
    st[2] = c->sym.flops_factor;
    st[3] = c->sym.max_front;
    st[4] = static_cast<double>(c->sym.leaf_ptr.size()) - 1;
    st[5] = static_cast<double>(c->sym.level_ptr.size()) - 1;
    st[6] = c->sym.g_entries;
    st[7] = c->fplan.n_levels;
    st[8] = static_cast<double>(c->fplan.launches.size());
    st[9] = c->fplan.flops;
    st[10] = static_cast<double>(c->fplan.pool_size);
    st[11] = c->sym.flops_g;
    return WRAPD_OK;
}

int wrapd_lbfgs_configure(wrapd_ctx* c, const wrapd_lbfgs_cfg* cfg) {
    CTX_OK(c);
    drop_begun(c);
    if (!cfg || cfg->memory < 1 || cfg->memory > kMaxHist || cfg->max_iters < 0) {
        if (c) c->last_error = "L-BFGS memory must be in [1, 8]";
        return WRAPD_ERR_ARG;
    }
    c->lcfg = *cfg;
    c->m_k = 0;
    return WRAPD_OK;
}

int wrapd_lbfgs_reset(wrapd_ctx* c) {
    CTX_OK(c);
    drop_begun(c);
    c->m_k = 0;
    return WRAPD_OK;
}

// The global step's first launches: value(x) fused with the first gradient (the gradient
// from the cache at the same x), and the first two-loop's alphas for the current history.
int wrapd_global_step_begin(wrapd_ctx* c) {
    CTX_OK(c); MESH_OK(c);
    if (!c->factored) { c->last_error = "A0 is required but wasn't supplied."; return WRAPD_ERR_STATE; }
    if (c->m_k == 0) launch_fill(c, c->grad_old, 0.0, 3ll * c->n_f);   // minimize() re-zeros its vectors when m_k == 0
    arm_tail(c, 1, 0u, S_VAL);
    launch_defo(c, DM_GRAD | DM_VALUE, 0);
    arm_tail(c, 1, 0u, S_DG);
    int hs[kMaxHist];
    const int h = std::min(c->lcfg.memory, c->m_k);
    for (int i = 0; i < h; ++i) hs[i] = c->pos2slot[i];
    launch_gather_grad(c, c->grad, 1, 0, nullptr, 0, hs, h);
    c->gs_begun = true;
    return launch_status(c);
}

// SplitLBFGS::minimize between the two GlobalProblem::value calls of RotAwareGS::global_step.
int wrapd_global_step(wrapd_ctx* c, double tol, wrapd_global_out* out) {
    CTX_OK(c); MESH_OK(c);
    if (!c->factored) { c->last_error = "A0 is required but wasn't supplied."; return WRAPD_ERR_STATE; }
    if (!c->gs_begun) {
        const int st = wrapd_global_step_begin(c);
        if (st != WRAPD_OK) return st;
    }
    c->gs_begun = false;
    const int M = c->lcfg.memory;
    const long long n3 = 3ll * c->n_f;
    const double m_tol = tol;
    const double kappa = c->lcfg.kappa;
    wrapd_global_out o{};
    o.exit_reason = 0;
    // An iteration's first half: grad_old = grad (a pointer swap; x_old is kept implicitly by
    // the x-step), the two-loop's solve and the decrement g . r, which the host reads next.
    // It is queued before the host reads what the launches before it produced (the first
    // gradient, or the previous iteration's y . s), so the GPU has work while the host
    // reads; when the host then stops the loop, it undoes the swap (the rest wrote scratch).
    struct Stage {
        const double* g = nullptr;
        int it = 0;
        int slots[kMaxHist] = {};
        bool lazy_r = false;   // r = rv + sum coef_j s_j is formed by the x-step
        int dir_at = S_DIR;    // where the host reads g . r
        bool live = false;     // queued and not consumed
    };
    auto stage = [&](Stage& st) {
        std::swap(c->grad, c->grad_old);
        st.g = c->grad_old;
        st.it = std::min(M, c->m_k);
        for (int i = 0; i < st.it; ++i) st.slots[i] = c->pos2slot[i];
        st.lazy_r = false;
        st.dir_at = S_DIR;
        if (st.it > 0) {
            double* vs[kMaxHist];
            for (int i = 0; i < st.it; ++i) vs[i] = c->hS[st.slots[i]];
            // the gather that produced g has left the alphas of this list in S_ALPHA
            const bool have_alpha = c->alpha_n == st.it && std::equal(st.slots, st.slots + st.it, c->alpha_slots);
            if (!have_alpha) launch_multidot(c, st.g, vs, st.it, 1, st.slots);   // and the alpha recurrence
            c->alpha_n = -1;
            launch_solve(c, st.g, c->rv, st.slots, st.it);   // of q = g - sum alpha_j y_j, formed by the solve
            for (int i = 0; i < st.it; ++i) vs[i] = c->hY[st.slots[i]];
            if (c->lcfg.linesearch == 1) {
                launch_multidot(c, c->rv, vs, st.it, 2, st.slots);   // and the beta recurrence
                launch_axpy_r(c, st.g, st.slots, st.it);
            } else {
                // without the line search, r is only needed by the x-step, which forms it;
                // g . r then comes from the beta launch
                launch_multidot(c, c->rv, vs, st.it, 3, st.slots, st.g);   // the beta recurrence and g . r
                st.lazy_r = true;
            }
        } else {
            launch_solve(c, st.g, c->rv);
            launch_dot(c, st.g, c->rv);
            st.dir_at = S_RED;
        }
        st.live = true;
    };
    auto unstage = [&](Stage& st) {
        if (!st.live) return;
        std::swap(c->grad, c->grad_old);
        st.live = false;
    };
    int max_iters = c->lcfg.max_iters;
    Stage st;
    if (max_iters >= 1) stage(st);
    const double* r = scal_read(c, S_DG, 2);   // S_DG, S_VAL (and the stage's decrement)
    if (!r) return WRAPD_ERR_HIP;
    const double dg = r[0];
    o.obj_bef = r[1];
    o.obj_aft = o.obj_bef;
    o.energy = c->last_energy;
    o.flips = c->last_flips;
    if (dg < 1.e-13) {
        unstage(st);
        o.iters = 0;
        o.exit_reason = 3;
        if (out) *out = o;
        return launch_status(c);
    }
    int global_iter = 1;
    bool breaking_out = false;
    bool have_final = false;
    double alpha_init = 1.0;
    for (; global_iter <= max_iters; ++global_iter) {
        if (!st.live) stage(st);
        const double* dp = scal_read(c, st.dir_at, 1);
        if (!dp) return WRAPD_ERR_HIP;
        const double dir = dp[0];
        st.live = false;
        const double* g = st.g;
        const int it = st.it;
        const int* slots = st.slots;
        bool lazy_r = st.lazy_r;
        if (std::isnan(dir)) {
            if (solve_aborted(c)) {
                c->last_error = "triangular solve: a dependency wait timed out";
                return WRAPD_ERR_HIP;
            }
            c->last_error = "Decrement is NaN.";
            return WRAPD_ERR_NAN;
        }
        const double* stepdir = c->rv;
        if (dir <= 0) {
            lazy_r = false;
            stepdir = g;   // q = grad
            max_iters -= c->m_k;
            c->m_k = 0;
            if (c->lcfg.linesearch == 1 && m_tol > 1.e-12) {
                launch_dot(c, g, g);
                const double* d = scal_read(c, S_RED, 2);
                if (!d) return WRAPD_ERR_HIP;
                alpha_init = std::min(1.0, 1.0 / d[1]);
            }
        }
        double rate = 1.0;
        if (m_tol > 1.e-12 && c->lcfg.linesearch == 1) {
            // Backtracking::search on n x 3 (deps/mcloptlib/include/MCL/Backtracking.hpp:71-98), p = -q
            launch_dot(c, stepdir, stepdir);
            const double* d = scal_read(c, S_RED, 1);
            if (!d) return WRAPD_ERR_HIP;
            const double pnorm = std::sqrt(d[0]);
            if (pnorm <= 2.220446049250313e-16) {
                rate = 1e-4;
            } else {
                launch_dot(c, g, stepdir);
                d = scal_read(c, S_RED, 1);
                if (!d) return WRAPD_ERR_HIP;
                const double gtp = -d[0];
                arm_tail(c, 1, 0u, S_RED);
                launch_defo(c, DM_VALUE, 0);
                d = scal_read(c, S_RED, 1);
                if (!d) return WRAPD_ERR_HIP;
                const double fx0 = d[0];
                launch_copy(c, c->bv, c->xall, n3);   // save x
                double alpha = alpha_init;
                int lsit = 0;
                for (; lsit < 100000; ++lsit) {
                    launch_armijo_x(c, -alpha, stepdir, c->bv);   // x = x_save - alpha q
                    arm_tail(c, 1, 0u, S_RED);
                    launch_defo(c, DM_VALUE, 0);
                    d = scal_read(c, S_RED, 1);
                    if (!d) return WRAPD_ERR_HIP;
                    if (d[0] <= fx0 + alpha * 1e-4 * gtp) break;
                    alpha *= 0.7;
                }
                launch_copy(c, c->xall, c->bv, n3);   // restore x
                rate = lsit >= 100000 ? -1.0 : alpha;
            }
            if (rate <= 0) { c->last_error = "Failure in linesearch"; return WRAPD_ERR_LINESEARCH; }
        }
        // position of the new (s, y) pair
        int new_pos, new_slot;
        if (c->m_k < M) {
            new_pos = c->m_k;
            new_slot = c->pos2slot[new_pos];
        } else {
            new_slot = c->pos2slot[0];
            for (int i = 0; i < M - 1; ++i) c->pos2slot[i] = c->pos2slot[i + 1];
            c->pos2slot[M - 1] = new_slot;
            new_pos = M - 1;
        }
        const int nlist = new_pos + 1;
        int lslots[kMaxHist];
        for (int i = 0; i < nlist; ++i) lslots[i] = c->pos2slot[i];
        if (lazy_r) launch_xstep_r(c, rate, slots, it, c->hS[new_slot]);
        else launch_xstep(c, rate, stepdir, c->hS[new_slot]);
        const bool early = dir > 0 && (dir < kappa * m_tol || dir < kappa * 1.e-12);
        const bool last = early || (global_iter + 1 > max_iters);
        arm_tail(c, last ? 3 : 1, 0u, S_VAL);
        launch_defo(c, last ? (DM_GRAD | DM_VALUE | DM_WRITE_SYMF | DM_ENERGY) : (DM_GRAD | DM_VALUE), 0);
        launch_gather_grad(c, c->grad, 2, new_pos, lslots, nlist, lslots, nlist);   // and the next alphas
        if (!last) {   // the loop goes on unless y . s is 0: the next iteration's first half goes out now
            c->m_k++;
            stage(st);
        }
        const double* d = scal_read(c, S_DENOM, 5);   // DENOM, DG, VAL, EN, FLIPS
        if (!d) return WRAPD_ERR_HIP;
        const double denom = d[0];
        o.obj_aft = d[2];
        if (last) {
            o.energy = d[3];
            o.flips = static_cast<int>(d[4]);
            have_final = true;
        }
        if (early) {
            c->m_k += 1;
            breaking_out = true;
            o.exit_reason = 1;
            break;
        }
        if (std::fabs(denom) <= 0) {
            if (st.live) {   // the pair is not kept: neither its count nor the next iteration's launches
                unstage(st);
                c->m_k--;
            }
            breaking_out = true;
            o.exit_reason = 2;
            have_final = false;
            break;
        }
        alpha_init = 1.0;
        if (!st.live) c->m_k++;   // (counted before the next iteration was queued otherwise)
    }
    c->m_k = std::min(M, c->m_k);
    o.iters = breaking_out ? global_iter : global_iter - 1;
    if (!have_final) {
        arm_tail(c, 3, 0u, S_RED);
        launch_defo(c, DM_VALUE | DM_WRITE_SYMF | DM_ENERGY, 0);
        const double* d = scal_read(c, S_RED, 3);
        if (!d) return WRAPD_ERR_HIP;
        o.obj_aft = d[0];
        o.energy = d[1];
        o.flips = static_cast<int>(d[2]);
    }
    c->last_energy = o.energy;
    c->last_flips = o.flips;
    if (out) *out = o;
    return launch_status(c);
}

// UnawareGS::global_step: b = add_to_b, x = A^-1 b, refresh the cache (F and sigma)
int wrapd_global_step_unaware(wrapd_ctx* c, wrapd_global_out* out) {
    CTX_OK(c); MESH_OK(c);
    drop_begun(c);
    if (!c->factored) { c->last_error = "no factorization"; return WRAPD_ERR_STATE; }
    launch_defo(c, DM_BVEC, 0);
    launch_gather_grad(c, c->bv, 0, 0, nullptr, 0);
    launch_solve(c, c->bv, c->xall);
    arm_tail(c, 3, 0u, S_RED);
    launch_defo(c, DM_FT_CACHE | DM_ENERGY, 0);
    const double* d = scal_read(c, S_RED, 3);
    if (!d) return WRAPD_ERR_HIP;
    if (std::isnan(d[1]) && solve_aborted(c)) {
        c->last_error = "triangular solve: a dependency wait timed out";
        return WRAPD_ERR_HIP;
    }
    wrapd_global_out o{};
    o.iters = 1;
    o.energy = d[1];
    o.flips = static_cast<int>(d[2]);
    c->last_energy = o.energy;
    c->last_flips = o.flips;
    if (out) *out = o;
    return launch_status(c);
}

int wrapd_add_to_b(wrapd_ctx* c, double* b) {
    CTX_OK(c); MESH_OK(c);
    if (!b) return WRAPD_ERR_ARG;
    drop_begun(c);
    launch_defo(c, DM_BVEC, 0);
    launch_gather_grad(c, c->bv, 0, 0, nullptr, 0);
    if (d2h(c, b, c->bv, 3ull * c->n_f)) return WRAPD_ERR_HIP;
    return launch_status(c);
}

int wrapd_get_elem(wrapd_ctx* c, const char* name, double* outp) {
    CTX_OK(c); MESH_OK(c);
    if (!name || !outp) return WRAPD_ERR_ARG;
    const std::string n(name);
    const int ne = c->n_e, nep = c->n_ep;
    auto soa = [&](const double* d, int ncomp) -> int {
        std::vector<double> h(static_cast<size_t>(ncomp) * nep);
        if (d2h(c, h.data(), d, h.size())) return WRAPD_ERR_HIP;
        for (int e = 0; e < ne; ++e)
            for (int q = 0; q < ncomp; ++q) outp[static_cast<size_t>(e) * ncomp + q] = h[static_cast<size_t>(q) * nep + e];
        return WRAPD_OK;
    };
    if (n == "U") return soa(c->U, 9);
    if (n == "P") return soa(c->P, 9);
    if (n == "symF") return soa(c->symF, 9);
    if (n == "wsq") return soa(c->wsq, 1);
    if (n == "cand_wsq") return soa(c->cand, 1);
    if (n == "sigma_last") return soa(c->sig, 3);
    if (n == "vol") return soa(c->vol, 1);
    if (n == "Di") return soa(c->Di, 12);
    if (n == "Z") {
        std::vector<double> u(9ull * ne);
        int st = soa(c->U, 9);
        if (st) return st;
        std::copy(outp, outp + 9ull * ne, u.begin());
        st = soa(c->P, 9);
        if (st) return st;
        for (size_t i = 0; i < u.size(); ++i) outp[i] += u[i];
        return WRAPD_OK;
    }
    c->last_error = "unknown element array";
    return WRAPD_ERR_ARG;
}

// ---- checkpoint / resume of the device-resident ADMM state ----------------------------
extern "C++" {
namespace {
struct StateHeader {
    char magic[8];   // "WRPDST01"
    int n_e, n_ep, n_f, n_fix, nhist, rot_aware, m_k, last_flips, grad_old_valid;
    int pos2slot[wd::kMaxHist];
    double last_energy;
    wrapd_lbfgs_cfg lcfg;
};
// (device pointer, doubles) of every array in the state, in blob order
std::vector<std::pair<double*, size_t>> state_arrays(wrapd_ctx* c) {
    const size_t ne = c->n_ep, nv = 3ull * c->n_f;
    std::vector<std::pair<double*, size_t>> a = {
        {c->wsq, ne}, {c->cand, ne}, {c->U, 9 * ne}, {c->P, 9 * ne}, {c->symF, 9 * ne}, {c->sig, 3 * ne},
        {c->xall, 3ull * (c->n_f + c->n_fix)}, {c->grad, nv}, {c->grad_old, nv}, {c->scal, static_cast<size_t>(S_NUM)}};
    for (int k = 0; k < kMaxHist; ++k) a.push_back({c->hS[k], nv});
    for (int k = 0; k < kMaxHist; ++k) a.push_back({c->hY[k], nv});
    return a;
}
}  // namespace
}  // extern "C++"

int wrapd_state_save(wrapd_ctx* c, void* buf, size_t cap, size_t* size) {
    CTX_OK(c); MESH_OK(c);
    if (!size) return WRAPD_ERR_ARG;
    const auto arrs = state_arrays(c);
    size_t total = sizeof(StateHeader);
    for (const auto& a : arrs) total += a.second * sizeof(double);
    *size = total;
    if (!buf) return WRAPD_OK;
    if (cap < total) { c->last_error = "wrapd_state_save: buffer too small"; return WRAPD_ERR_ARG; }
    StateHeader h{};
    std::memcpy(h.magic, "WRPDST01", 8);
    h.n_e = c->n_e; h.n_ep = c->n_ep; h.n_f = c->n_f; h.n_fix = c->n_fix; h.nhist = kMaxHist;
    h.rot_aware = c->rot_aware; h.m_k = c->m_k; h.last_flips = c->last_flips; h.grad_old_valid = c->grad_old_valid;
    for (int k = 0; k < kMaxHist; ++k) h.pos2slot[k] = c->pos2slot[k];
    h.last_energy = c->last_energy;
    h.lcfg = c->lcfg;
    char* p = static_cast<char*>(buf);
    std::memcpy(p, &h, sizeof(h));
    p += sizeof(h);
    for (const auto& a : arrs) {
        if (check_hip(c, hipMemcpyAsync(p, a.first, a.second * sizeof(double), hipMemcpyDeviceToHost, c->stream),
                      "state D2H"))
            return WRAPD_ERR_HIP;
        p += a.second * sizeof(double);
    }
    return check_hip(c, hipStreamSynchronize(c->stream), "sync") ? WRAPD_ERR_HIP : WRAPD_OK;
}

int wrapd_state_load(wrapd_ctx* c, const void* buf, size_t size) {
    CTX_OK(c); MESH_OK(c);
    drop_begun(c);
    if (!buf || size < sizeof(StateHeader)) return WRAPD_ERR_ARG;
    StateHeader h;
    std::memcpy(&h, buf, sizeof(h));
    if (std::memcmp(h.magic, "WRPDST01", 8) != 0) { c->last_error = "wrapd_state_load: not a state blob"; return WRAPD_ERR_ARG; }
    if (h.n_e != c->n_e || h.n_ep != c->n_ep || h.n_f != c->n_f || h.n_fix != c->n_fix || h.nhist != kMaxHist) {
        c->last_error = "wrapd_state_load: the state belongs to another mesh";
        return WRAPD_ERR_ARG;
    }
    const auto arrs = state_arrays(c);
    size_t total = sizeof(StateHeader);
    for (const auto& a : arrs) total += a.second * sizeof(double);
    if (size != total) { c->last_error = "wrapd_state_load: size mismatch"; return WRAPD_ERR_ARG; }
    {
        // the L-BFGS bookkeeping indexes device buffers: reject what no save can produce
        bool seen[kMaxHist] = {};
        bool perm = true;
        for (int k = 0; k < kMaxHist; ++k) {
            const int v = h.pos2slot[k];
            if (v < 0 || v >= kMaxHist || seen[v]) { perm = false; break; }
            seen[v] = true;
        }
        if (!perm || h.lcfg.memory < 1 || h.lcfg.memory > kMaxHist || h.m_k < 0 || h.m_k > h.lcfg.memory ||
            (h.rot_aware != 0 && h.rot_aware != 1)) {
            c->last_error = "wrapd_state_load: corrupt L-BFGS state";
            return WRAPD_ERR_ARG;
        }
    }
    const char* p = static_cast<const char*>(buf) + sizeof(h);
    for (const auto& a : arrs) {
        if (check_hip(c, hipMemcpyAsync(a.first, p, a.second * sizeof(double), hipMemcpyHostToDevice, c->stream),
                      "state H2D"))
            return WRAPD_ERR_HIP;
        if (a.first == c->scal) std::memcpy(c->scal_host, p, a.second * sizeof(double));   // the host's copy too
        p += a.second * sizeof(double);
    }
    if (check_hip(c, hipStreamSynchronize(c->stream), "sync")) return WRAPD_ERR_HIP;
    c->rot_aware = h.rot_aware;
    c->m_k = h.m_k;   // checked below, before any device copy
    c->last_flips = h.last_flips;
    c->grad_old_valid = h.grad_old_valid != 0;
    for (int k = 0; k < kMaxHist; ++k) c->pos2slot[k] = h.pos2slot[k];
    c->last_energy = h.last_energy;
    c->lcfg = h.lcfg;
    c->factored = false;   // the factorization depends on w^2: wrapd_refactor before the next solve
    return WRAPD_OK;
}

int wrapd_svd3_batch(wrapd_ctx* c, const double* F, int n, int precision, double* sigma, double* U, double* V) {
    CTX_OK(c);
    if (n < 0 || (n > 0 && (!F || !sigma || !U || !V)) || (precision != WRAPD_FP64 && precision != WRAPD_FP32))
        return WRAPD_ERR_ARG;
    if (n == 0) return WRAPD_OK;
    double* d = nullptr;
    int st = WRAPD_OK;
    if (dalloc(c, d, 30ull * n)) return WRAPD_ERR_HIP;
    if (h2d(c, d, F, 9ull * n)) st = WRAPD_ERR_HIP;
    if (!st) {
        launch_svd_batch(c, n, precision, d, d + 9ll * n, d + 12ll * n, d + 21ll * n);
        if (launch_status(c) || d2h(c, sigma, d + 9ll * n, 3ull * n) || d2h(c, U, d + 12ll * n, 9ull * n) ||
            d2h(c, V, d + 21ll * n, 9ull * n))
            st = WRAPD_ERR_HIP;
    }
    (void)hipFree(d);
    return st;
}

int wrapd_set_elem(wrapd_ctx* c, const char* name, const double* in) {
    CTX_OK(c); MESH_OK(c);
    drop_begun(c);
    if (!name || !in) return WRAPD_ERR_ARG;
    const std::string n(name);
    const int ne = c->n_e, nep = c->n_ep;
    auto soa = [&](double* d, int ncomp) -> int {
        std::vector<double> h(static_cast<size_t>(ncomp) * nep, 0.0);
        if (d2h(c, h.data(), d, h.size())) return WRAPD_ERR_HIP;
        for (int e = 0; e < ne; ++e)
            for (int q = 0; q < ncomp; ++q) h[static_cast<size_t>(q) * nep + e] = in[static_cast<size_t>(e) * ncomp + q];
        if (h2d(c, d, h.data(), h.size())) return WRAPD_ERR_HIP;
        return check_hip(c, hipStreamSynchronize(c->stream), "sync") ? WRAPD_ERR_HIP : WRAPD_OK;
    };
    if (n == "U") return soa(c->U, 9);
    if (n == "P") return soa(c->P, 9);
    if (n == "symF") return soa(c->symF, 9);
    if (n == "wsq") return soa(c->wsq, 1);
    if (n == "cand_wsq") return soa(c->cand, 1);
    if (n == "sigma_last") return soa(c->sig, 3);
    c->last_error = "unknown or read-only element array";
    return WRAPD_ERR_ARG;
}

int wrapd_profile_enable(wrapd_ctx* c, int on) {
    CTX_OK(c);
    prof_flush(c);
    c->prof.on = on != 0;
    if (on) c->prof.acc.clear();
    return WRAPD_OK;
}

int wrapd_profile_only(wrapd_ctx* c, const char* only) {
    CTX_OK(c);
    prof_flush(c);
    c->prof.only = only ? only : "";
    return WRAPD_OK;
}

int wrapd_profile_get(wrapd_ctx* c, const char* kernel, double* total_ms, double* launches) {
    CTX_OK(c);
    prof_flush(c);
    auto it = c->prof.acc.find(kernel ? kernel : "");
    if (it == c->prof.acc.end()) {
        if (total_ms) *total_ms = 0.0;
        if (launches) *launches = 0.0;
        return WRAPD_OK;
    }
    if (total_ms) *total_ms = it->second.first;
    if (launches) *launches = it->second.second;
    return WRAPD_OK;
}

int wrapd_synchronize(wrapd_ctx* c) {
    CTX_OK(c);
    return check_hip(c, hipStreamSynchronize(c->stream), "sync") ? WRAPD_ERR_HIP : WRAPD_OK;
}

}  // extern "C"

// ---- learned prox warm start -------------------------------------------------------
namespace {
std::vector<float> mlp_download_sample(wrapd_ctx* c, long long max_n) {
    const long long n = c->mlp_log_count;
    const long long stride = std::max<long long>(1, n / std::max<long long>(1, max_n));
    const long long m = (n + stride - 1) / stride;
    std::vector<float> out(static_cast<size_t>(m) * 9);
    if (m == 0) return out;
    if (mlp_sample(c, stride, m, out.data()) != WRAPD_OK) out.clear();
    return out;
}
}  // namespace

extern "C" {

int wrapd_mlp_set_params(wrapd_ctx* c, const float* params, const float* norm) {
    CTX_OK(c);
    if (!params || !norm) return WRAPD_ERR_ARG;
    for (int i = 6; i < 12; ++i)
        if (!(norm[i] > 0.0f)) { c->last_error = "mlp: input standard deviations must be positive"; return WRAPD_ERR_ARG; }
    for (int i = 15; i < 18; ++i)
        if (!(norm[i] > 0.0f)) { c->last_error = "mlp: output standard deviations must be positive"; return WRAPD_ERR_ARG; }
    int st = mlp_upload(c, params, norm);
    if (st) return st;
    if (check_hip(c, hipMemsetAsync(c->mlp_m1, 0, kMlpParams * sizeof(float), c->stream), "mlp") ||
        check_hip(c, hipMemsetAsync(c->mlp_m2, 0, kMlpParams * sizeof(float), c->stream), "mlp"))
        return WRAPD_ERR_HIP;
    c->mlp_t = 0;
    return launch_status(c);
}

int wrapd_mlp_get_params(wrapd_ctx* c, float* params, float* norm) {
    CTX_OK(c);
    if (!c->mlp_ready) { c->last_error = "mlp: no network"; return WRAPD_ERR_STATE; }
    if ((params && check_hip(c, hipMemcpyAsync(params, c->mlp_master, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, c->stream), "mlp")) ||
        (norm && check_hip(c, hipMemcpyAsync(norm, c->mlp_norm, kMlpNorm * sizeof(float), hipMemcpyDeviceToHost, c->stream), "mlp")))
        return WRAPD_ERR_HIP;
    return check_hip(c, hipStreamSynchronize(c->stream), "mlp") ? WRAPD_ERR_HIP : WRAPD_OK;
}

int wrapd_mlp_init(wrapd_ctx* c, unsigned long long seed) {
    CTX_OK(c);
    // Glorot-uniform weights, zero biases, identity normalisation
    std::vector<float> p(kMlpParams, 0.0f), norm(kMlpNorm, 0.0f);
    uint64_t z = seed;
    auto next = [&z]() {
        z += 0x9E3779B97F4A7C15ull;
        uint64_t r = z;
        r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ull;
        r = (r ^ (r >> 27)) * 0x94D049BB133111EBull;
        r ^= r >> 31;
        return static_cast<double>(r >> 11) * (1.0 / 9007199254740992.0);
    };
    auto fill = [&](int off, int rows, int cols) {
        const double a = std::sqrt(6.0 / (rows + cols));
        for (int i = 0; i < rows * cols; ++i) p[off + i] = static_cast<float>((2.0 * next() - 1.0) * a);
    };
    fill(kMlpOffW1, kMlpHid, kMlpIn);
    fill(kMlpOffW2, kMlpHid, kMlpHid);
    fill(kMlpOffW3, kMlpOut, kMlpHid);
    for (int i = 6; i < 12; ++i) norm[i] = 1.0f;
    for (int i = 15; i < 18; ++i) norm[i] = 1.0f;
    return wrapd_mlp_set_params(c, p.data(), norm.data());
}

int wrapd_mlp_infer(wrapd_ctx* c, const float* in, int n, float* out) {
    CTX_OK(c);
    if (!c->mlp_ready) { c->last_error = "mlp: no network"; return WRAPD_ERR_STATE; }
    if (n < 0 || (n > 0 && (!in || !out))) return WRAPD_ERR_ARG;
    if (n == 0) return WRAPD_OK;
    float *din = nullptr, *dout = nullptr;
    int st = WRAPD_OK;
    if (check_hip(c, hipMalloc(&din, 6ull * n * sizeof(float)), "mlp infer") ||
        check_hip(c, hipMalloc(&dout, 3ull * n * sizeof(float)), "mlp infer") ||
        check_hip(c, hipMemcpyAsync(din, in, 6ull * n * sizeof(float), hipMemcpyHostToDevice, c->stream), "mlp infer")) {
        st = WRAPD_ERR_HIP;
    } else {
        mlp_infer(c, din, n, dout);
        if (check_hip(c, hipMemcpyAsync(out, dout, 3ull * n * sizeof(float), hipMemcpyDeviceToHost, c->stream), "mlp infer") ||
            check_hip(c, hipStreamSynchronize(c->stream), "mlp infer"))
            st = WRAPD_ERR_HIP;
    }
    if (din) (void)hipFree(din);
    if (dout) (void)hipFree(dout);
    return st ? st : launch_status(c);
}

int wrapd_mlp_grad(wrapd_ctx* c, const float* tuples, int n, float* grad, float* loss) {
    CTX_OK(c);
    if (!c->mlp_ready) { c->last_error = "mlp: no network"; return WRAPD_ERR_STATE; }
    if (n <= 0 || !tuples) return WRAPD_ERR_ARG;
    float* dt = nullptr;
    int* didx = nullptr;
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    int st = WRAPD_OK;
    if (check_hip(c, hipMalloc(&dt, 9ull * n * sizeof(float)), "mlp grad") ||
        check_hip(c, hipMalloc(&didx, static_cast<size_t>(n) * sizeof(int)), "mlp grad") ||
        check_hip(c, hipMemcpyAsync(dt, tuples, 9ull * n * sizeof(float), hipMemcpyHostToDevice, c->stream), "mlp grad") ||
        check_hip(c, hipMemcpyAsync(didx, idx.data(), static_cast<size_t>(n) * sizeof(int), hipMemcpyHostToDevice, c->stream), "mlp grad")) {
        st = WRAPD_ERR_HIP;
    } else {
        st = mlp_step(c, dt, n, didx, n, 0, 0.0f, 1);
        if (!st && ((grad && check_hip(c, hipMemcpyAsync(grad, c->mlp_gsum, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, c->stream), "mlp grad")) ||
                    (loss && check_hip(c, hipMemcpyAsync(loss, c->mlp_loss, sizeof(float), hipMemcpyDeviceToHost, c->stream), "mlp grad")) ||
                    check_hip(c, hipStreamSynchronize(c->stream), "mlp grad")))
            st = WRAPD_ERR_HIP;
    }
    if (dt) (void)hipFree(dt);
    if (didx) (void)hipFree(didx);
    return st ? st : launch_status(c);
}

int wrapd_mlp_grad_noalloc(wrapd_ctx* c, const float* tuples, int n, float* grad, float* loss) {
    CTX_OK(c);
    if (!c->mlp_ready) { c->last_error = "mlp: no network"; return WRAPD_ERR_STATE; }
    if (n <= 0 || !tuples) return WRAPD_ERR_ARG;
    if (n > kMlpGradNoAllocMax) {
        c->last_error = "wrapd_mlp_grad_noalloc: at most 65536 tuples per call (the static scratch array is 8 MiB)";
        return WRAPD_ERR_ARG;
    }
    if (!c->scratch) { c->last_error = "wrapd_mlp_grad_noalloc: no scratch address for this device"; return WRAPD_ERR_STATE; }
    const MlpGradScratch L = mlp_grad_layout(n);
    unsigned char* base = static_cast<unsigned char*>(c->scratch);
    float* d_partial = reinterpret_cast<float*>(base + L.partial);
    float* d_lpart = reinterpret_cast<float*>(base + L.lpart);
    float* d_gsum = reinterpret_cast<float*>(base + L.gsum);
    float* d_loss = reinterpret_cast<float*>(base + L.loss);
    float* d_tuples = reinterpret_cast<float*>(base + L.tuples);
    int* d_idx = reinterpret_cast<int*>(base + L.idx);
    // the sample list 0..n-1, as wrapd_mlp_grad uses it; a small host vector, no device allocation
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    if (check_hip(c, hipMemcpyAsync(d_tuples, tuples, 9ull * n * sizeof(float), hipMemcpyHostToDevice, c->stream),
                  "mlp grad (no alloc)") ||
        check_hip(c, hipMemcpyAsync(d_idx, idx.data(), static_cast<size_t>(n) * sizeof(int), hipMemcpyHostToDevice,
                                    c->stream), "mlp grad (no alloc)"))
        return WRAPD_ERR_HIP;
    launch_mlp_grad(c, d_tuples, d_idx, n, d_partial, d_lpart, d_gsum, d_loss);
    if ((grad && check_hip(c, hipMemcpyAsync(grad, d_gsum, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, c->stream),
                           "mlp grad (no alloc)")) ||
        (loss && check_hip(c, hipMemcpyAsync(loss, d_loss, sizeof(float), hipMemcpyDeviceToHost, c->stream),
                           "mlp grad (no alloc)")) ||
        check_hip(c, hipStreamSynchronize(c->stream), "mlp grad (no alloc)"))
        return WRAPD_ERR_HIP;
    return launch_status(c);
}

int wrapd_mlp_log_start(wrapd_ctx* c, long long capacity) {
    CTX_OK(c);
    if (capacity < 0) return WRAPD_ERR_ARG;
    if (capacity > c->mlp_log_cap) {
        if (c->mlp_log) (void)hipFree(c->mlp_log);
        c->mlp_log = nullptr;
        c->mlp_log_cap = 0;
        if (check_hip(c, hipMalloc(&c->mlp_log, std::max<long long>(1, capacity) * 9 * sizeof(float)), "mlp log"))
            return WRAPD_ERR_HIP;
        c->mlp_log_cap = capacity;
    }
    c->mlp_log_count = 0;
    c->mlp_logging = true;
    return WRAPD_OK;
}

int wrapd_mlp_log_stop(wrapd_ctx* c) {
    CTX_OK(c);
    c->mlp_logging = false;
    return WRAPD_OK;
}

long long wrapd_mlp_log_count(wrapd_ctx* c) { return c ? c->mlp_log_count : -1; }

int wrapd_mlp_log_get(wrapd_ctx* c, long long offset, long long n, float* out) {
    CTX_OK(c);
    if (offset < 0 || n < 0 || offset + n > c->mlp_log_count || (n > 0 && !out)) return WRAPD_ERR_ARG;
    if (n == 0) return WRAPD_OK;
    if (check_hip(c, hipMemcpyAsync(out, c->mlp_log + 9 * offset, 9 * n * sizeof(float), hipMemcpyDeviceToHost, c->stream), "mlp log") ||
        check_hip(c, hipStreamSynchronize(c->stream), "mlp log"))
        return WRAPD_ERR_HIP;
    return WRAPD_OK;
}

int wrapd_mlp_log_set(wrapd_ctx* c, const float* tuples, long long n) {
    CTX_OK(c);
    if (n < 0 || (n > 0 && !tuples)) return WRAPD_ERR_ARG;
    int st = wrapd_mlp_log_start(c, n);
    c->mlp_logging = false;   // a set log is data, not a recording
    if (st) return st;
    if (n > 0 && check_hip(c, hipMemcpyAsync(c->mlp_log, tuples, 9 * n * sizeof(float), hipMemcpyHostToDevice, c->stream), "mlp log"))
        return WRAPD_ERR_HIP;
    c->mlp_log_count = n;
    return launch_status(c);
}

int wrapd_mlp_fit_norm(wrapd_ctx* c) {
    CTX_OK(c);
    if (!c->mlp_ready) { c->last_error = "mlp: no network"; return WRAPD_ERR_STATE; }
    if (c->mlp_log_count < 2) { c->last_error = "mlp: no logged tuples"; return WRAPD_ERR_STATE; }
    const std::vector<float> t = mlp_download_sample(c, 1000000);
    if (t.empty()) { c->last_error = "mlp: log download failed"; return WRAPD_ERR_HIP; }
    const size_t m = t.size() / 9;
    double s[9] = {0}, q[9] = {0};
    for (size_t i = 0; i < m; ++i) {
        const float* r = &t[9 * i];
        const double v[9] = {r[0], r[1], r[2], std::log(r[3] / r[4]), std::log(r[5] / r[4]), std::log(r[4]),
                             r[6] - r[0], r[7] - r[1], r[8] - r[2]};
        for (int k = 0; k < 9; ++k) { s[k] += v[k]; q[k] += v[k] * v[k]; }
    }
    std::vector<float> norm(kMlpNorm);
    for (int k = 0; k < 9; ++k) {
        const double mean = s[k] / m;
        const double sd = std::sqrt(std::max(0.0, q[k] / m - mean * mean));
        const double sdc = sd > 1e-6 * (1.0 + std::fabs(mean)) ? sd : 1.0;   // constant features: unit scale
        if (k < 6) { norm[k] = static_cast<float>(mean); norm[6 + k] = static_cast<float>(sdc); }
        else { norm[12 + k - 6] = static_cast<float>(mean); norm[15 + k - 6] = static_cast<float>(sdc); }
    }
    if (check_hip(c, hipMemcpyAsync(c->mlp_norm, norm.data(), kMlpNorm * sizeof(float), hipMemcpyHostToDevice, c->stream), "mlp norm") ||
        check_hip(c, hipStreamSynchronize(c->stream), "mlp norm"))
        return WRAPD_ERR_HIP;
    return WRAPD_OK;
}

int wrapd_mlp_augment(wrapd_ctx* c, double w_lo, double w_hi, unsigned long long seed) {
    CTX_OK(c);
    if (c->mlp_log_count < 1) { c->last_error = "mlp: no logged tuples"; return WRAPD_ERR_STATE; }
    if (!(w_lo > 0.0) || !(w_hi > 0.0)) {   // the logged range of w^2/vol
        const std::vector<float> t = mlp_download_sample(c, 1000000);
        if (t.empty()) return WRAPD_ERR_HIP;
        double lo = 1e300, hi = 0.0;
        for (size_t i = 0; i < t.size() / 9; ++i) { lo = std::min(lo, double(t[9 * i + 3])); hi = std::max(hi, double(t[9 * i + 3])); }
        w_lo = lo;
        w_hi = hi;
    }
    if (!(w_hi >= w_lo)) return WRAPD_ERR_ARG;
    launch_mlp_augment(c, c->mlp_log, c->mlp_log_count, w_lo, w_hi, seed);
    return launch_status(c);
}

int wrapd_mlp_train_full(wrapd_ctx* c, const float* tuples, int n, int steps, float lr, float* loss_hist) {
    CTX_OK(c);
    if (!c->mlp_ready) { c->last_error = "mlp: no network"; return WRAPD_ERR_STATE; }
    if (steps < 0 || n <= 0 || !tuples || !(lr > 0.0f)) return WRAPD_ERR_ARG;
    float* dt = nullptr;
    int* didx = nullptr;
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    int st = WRAPD_OK;
    if (check_hip(c, hipMalloc(&dt, 9ull * n * sizeof(float)), "mlp train") ||
        check_hip(c, hipMalloc(&didx, static_cast<size_t>(n) * sizeof(int)), "mlp train") ||
        check_hip(c, hipMemcpyAsync(dt, tuples, 9ull * n * sizeof(float), hipMemcpyHostToDevice, c->stream), "mlp train") ||
        check_hip(c, hipMemcpyAsync(didx, idx.data(), static_cast<size_t>(n) * sizeof(int), hipMemcpyHostToDevice,
                                    c->stream), "mlp train"))
        st = WRAPD_ERR_HIP;
    for (int k = 0; k < steps && !st; ++k) {
        st = mlp_step(c, dt, n, didx, n, 0, lr, ++c->mlp_t);
        if (!st && loss_hist &&
            (check_hip(c, hipMemcpyAsync(loss_hist + k, c->mlp_loss, sizeof(float), hipMemcpyDeviceToHost, c->stream),
                       "mlp train") ||
             check_hip(c, hipStreamSynchronize(c->stream), "mlp train")))
            st = WRAPD_ERR_HIP;
    }
    if (!st) st = check_hip(c, hipStreamSynchronize(c->stream), "mlp train") ? WRAPD_ERR_HIP : WRAPD_OK;
    if (dt) (void)hipFree(dt);
    if (didx) (void)hipFree(didx);
    return st ? st : launch_status(c);
}

int wrapd_mlp_train(wrapd_ctx* c, int steps, int batch, float lr, unsigned long long seed, float* loss_hist) {
    CTX_OK(c);
    if (!c->mlp_ready) { c->last_error = "mlp: no network"; return WRAPD_ERR_STATE; }
    if (steps < 0 || batch <= 0 || !(lr > 0.0f)) return WRAPD_ERR_ARG;
    if (c->mlp_log_count < 1) { c->last_error = "mlp: no logged tuples"; return WRAPD_ERR_STATE; }
    float* dl = nullptr;
    if (loss_hist && check_hip(c, hipMalloc(&dl, std::max(1, steps) * sizeof(float)), "mlp train")) return WRAPD_ERR_HIP;
    int st = WRAPD_OK;
    {
        ProfScope ps(c, "mlp_train");
        for (int k = 0; k < steps && !st; ++k) {
            st = mlp_step(c, c->mlp_log, c->mlp_log_count, nullptr, batch, seed * 0x9E3779B97F4A7C15ull + k, lr, ++c->mlp_t);
            if (!st && dl) st = check_hip(c, hipMemcpyAsync(dl + k, c->mlp_loss, sizeof(float), hipMemcpyDeviceToDevice, c->stream), "mlp train") ? WRAPD_ERR_HIP : WRAPD_OK;
        }
    }
    if (!st && dl && steps > 0 &&
        (check_hip(c, hipMemcpyAsync(loss_hist, dl, steps * sizeof(float), hipMemcpyDeviceToHost, c->stream), "mlp train") ||
         check_hip(c, hipStreamSynchronize(c->stream), "mlp train")))
        st = WRAPD_ERR_HIP;
    if (dl) (void)hipFree(dl);
    return st ? st : launch_status(c);
}

}  // extern "C"
