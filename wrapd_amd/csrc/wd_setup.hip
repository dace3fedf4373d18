// Mesh maps built on the device: what TetElements::get_A_triplets / update_A_coeffs and the
// vertex gathers of the reference loop over on the host (src/TetElements.cpp:483-510,
// src/System.cpp:75-104).
//
//   vidx     element corner -> combined position index (free: [0, n_f), fixed: n_f + j)
//   A        the upper triangle of A = sum_e w_e^2 D_e^T D_e over the free vertices: its
//            pattern (row pointers and columns, also copied to the host for the symbolic
//            analysis) and, per entry, the list of the element contributions e*16 + a*4 + b
//            that sum into it, in element order (the reference's triplet order)
//   v2e_ptr  free vertex -> its element corners (4e + q), in element order; cpos is the
//            inverse (each corner's row of the gradient staging, -1 for a fixed vertex)
//
// Both maps are stable radix sorts of keys emitted in element order: A's keys are
// row * n_f + col for every corner pair (a, b) of an element with free vertices and
// row <= col, the vertex lists' keys the free vertex of each corner.  Equal keys keep the
// emission order, so each entry's contributions come out in the host loop's order and A's
// values sum in the same order as before.
#include <hipcub/hipcub.hpp>

#include "wd_ctx.hpp"

namespace wd {
namespace {

constexpr int kMapThreads = 256;

__device__ __forceinline__ int a_pairs(const int f[4], int n_f) {
    int n = 0;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) n += (f[a] < n_f && f[b] < n_f && f[a] <= f[b]) ? 1 : 0;
    return n;
}

// per element: vidx, and the number of A keys and free corners it emits
__global__ void __launch_bounds__(kMapThreads)
k_map_count(int n_e, int n_ep, int n_f, const int4* __restrict__ tets, const int* __restrict__ comb,
            int4* __restrict__ vidx, int* __restrict__ cnt_a, int* __restrict__ cnt_v) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_ep) return;
    if (e >= n_e) {
        vidx[e] = make_int4(0, 0, 0, 0);
        return;
    }
    const int4 t = tets[e];
    const int f[4] = {comb[t.x], comb[t.y], comb[t.z], comb[t.w]};
    vidx[e] = make_int4(f[0], f[1], f[2], f[3]);
    cnt_a[e] = a_pairs(f, n_f);
    cnt_v[e] = (f[0] < n_f) + (f[1] < n_f) + (f[2] < n_f) + (f[3] < n_f);
}

__global__ void __launch_bounds__(kMapThreads)
k_map_emit(int n_e, int n_f, const int4* __restrict__ vidx, const int* __restrict__ off_a,
           const int* __restrict__ off_v, unsigned long long* __restrict__ key_a, int* __restrict__ val_a,
           unsigned* __restrict__ key_v, int* __restrict__ val_v) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_e) return;
    const int4 t = vidx[e];
    const int f[4] = {t.x, t.y, t.z, t.w};
    int pa = off_a[e];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if (f[a] < n_f && f[b] < n_f && f[a] <= f[b]) {
                key_a[pa] = static_cast<unsigned long long>(f[a]) * static_cast<unsigned long long>(n_f) + f[b];
                val_a[pa] = e * 16 + a * 4 + b;
                ++pa;
            }
    int pv = off_v[e];
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (f[q] < n_f) {
            key_v[pv] = static_cast<unsigned>(f[q]);
            val_v[pv] = 4 * e + q;
            ++pv;
        }
}

// head[i] = 1 where sorted key i starts a run
__global__ void __launch_bounds__(kMapThreads)
k_map_heads(long long n, const unsigned long long* __restrict__ key, int* __restrict__ head) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    head[i] = (i == 0 || key[i] != key[i - 1]) ? 1 : 0;
}

// for each run (slot s = its inclusive head count - 1): its first contribution, row, column
__global__ void __launch_bounds__(kMapThreads)
k_map_slots(long long n, int n_f, const unsigned long long* __restrict__ key, const int* __restrict__ head,
            const int* __restrict__ slot_incl, int* __restrict__ cptr, int* __restrict__ srow, int* __restrict__ scol) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n || !head[i]) return;
    const int s = slot_incl[i] - 1;
    cptr[s] = static_cast<int>(i);
    srow[s] = static_cast<int>(key[i] / static_cast<unsigned long long>(n_f));
    scol[s] = static_cast<int>(key[i] % static_cast<unsigned long long>(n_f));
}

// row pointers from a row index per sorted item: ptr[r] = the first item of row >= r, for
// r in [0, n_rows]; item i covers the rows (row[i-1], row[i]] and item n the rows after the last
__global__ void __launch_bounds__(kMapThreads)
k_map_rowptr(long long n, int n_rows, const int* __restrict__ row, int* __restrict__ ptr) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i > n) return;
    const int r = i < n ? row[i] : n_rows;
    const int rp = i == 0 ? -1 : row[i - 1];
    for (int rr = rp + 1; rr <= r; ++rr) ptr[rr] = static_cast<int>(i);
}

__global__ void __launch_bounds__(kMapThreads)
k_map_cpos(long long n, const int* __restrict__ vlist, int* __restrict__ cpos) {
    const long long q = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= n) return;
    cpos[vlist[q]] = static_cast<int>(q);
}

__global__ void __launch_bounds__(kMapThreads)
k_map_fill(long long n, int v, int* __restrict__ p) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

int bits_for(unsigned long long n) {   // bits of n - 1 (keys lie in [0, n))
    int b = 0;
    while (b < 64 && (n - 1) >> b) ++b;
    return std::max(b, 1);
}

// a temporary device buffer, freed when it leaves scope
template <typename T>
struct DevBuf {
    T* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
    int alloc(wrapd_ctx* c, size_t n) { return check_hip(c, hipMalloc(&p, std::max<size_t>(1, n) * sizeof(T)), "hipMalloc"); }
};

}  // namespace

// d_tets: 4 n_ep vertex indices (padding rows arbitrary but in range), d_comb: per vertex
// its combined position index.  Allocates and fills c->vidx, c->a_cptr, c->a_cidx,
// c->v2e_ptr, c->cpos and sets c->nnzA, c->Ap, c->Aj; returns the number of free corners
// (the gradient staging's rows) in *n_corners.
int build_mesh_maps(wrapd_ctx* c, const int* d_tets, const int* d_comb, long long* n_corners) {
    const int n_e = c->n_e, n_ep = c->n_ep, n_f = c->n_f;
    hipStream_t s = c->stream;
    const int gb_e = blocks_for(n_ep, kMapThreads);
    auto alloc_out = [&](int*& p, size_t n) {
        if (p) { (void)hipFree(p); p = nullptr; }
        return check_hip(c, hipMalloc(&p, std::max<size_t>(1, n) * sizeof(int)), "hipMalloc");
    };
    if (alloc_out(c->vidx, 4ull * n_ep)) return WRAPD_ERR_HIP;
    DevBuf<int> cnt_a, cnt_v, off_a, off_v;
    if (cnt_a.alloc(c, n_e + 1) || cnt_v.alloc(c, n_e + 1) || off_a.alloc(c, n_e + 1) || off_v.alloc(c, n_e + 1))
        return WRAPD_ERR_HIP;
    hipLaunchKernelGGL(k_map_count, dim3(gb_e), dim3(kMapThreads), 0, s, n_e, n_ep, n_f,
                       reinterpret_cast<const int4*>(d_tets), d_comb, reinterpret_cast<int4*>(c->vidx), cnt_a.p, cnt_v.p);
    // exclusive offsets over n_e + 1 entries: the last is the total (the count past the end is 0)
    hipLaunchKernelGGL(k_map_fill, dim3(1), dim3(kMapThreads), 0, s, 1ll, 0, cnt_a.p + n_e);
    hipLaunchKernelGGL(k_map_fill, dim3(1), dim3(kMapThreads), 0, s, 1ll, 0, cnt_v.p + n_e);
    size_t tb_scan = 0;
    if (check_hip(c, hipcub::DeviceScan::ExclusiveSum(nullptr, tb_scan, cnt_a.p, off_a.p, n_e + 1, s), "scan size"))
        return WRAPD_ERR_HIP;
    DevBuf<unsigned char> tmp_scan;
    if (tmp_scan.alloc(c, tb_scan)) return WRAPD_ERR_HIP;
    if (check_hip(c, hipcub::DeviceScan::ExclusiveSum(tmp_scan.p, tb_scan, cnt_a.p, off_a.p, n_e + 1, s), "scan") ||
        check_hip(c, hipcub::DeviceScan::ExclusiveSum(tmp_scan.p, tb_scan, cnt_v.p, off_v.p, n_e + 1, s), "scan"))
        return WRAPD_ERR_HIP;
    int tot[2] = {0, 0};
    if (check_hip(c, hipMemcpyAsync(&tot[0], off_a.p + n_e, sizeof(int), hipMemcpyDeviceToHost, s), "D2H") ||
        check_hip(c, hipMemcpyAsync(&tot[1], off_v.p + n_e, sizeof(int), hipMemcpyDeviceToHost, s), "D2H") ||
        check_hip(c, hipStreamSynchronize(s), "sync"))
        return WRAPD_ERR_HIP;
    const long long na = tot[0], nv = tot[1];
    DevBuf<unsigned long long> key_a, key_a2;
    DevBuf<int> val_a, head, slot_incl, srow;
    DevBuf<unsigned> key_v, key_v2;
    DevBuf<int> val_v, vlist;
    if (key_a.alloc(c, na) || key_a2.alloc(c, na) || val_a.alloc(c, na) || key_v.alloc(c, nv) || key_v2.alloc(c, nv) ||
        val_v.alloc(c, nv) || vlist.alloc(c, nv))
        return WRAPD_ERR_HIP;
    if (alloc_out(c->a_cidx, na) || alloc_out(c->v2e_ptr, n_f + 1ull) || alloc_out(c->cpos, 4ull * n_ep))
        return WRAPD_ERR_HIP;
    hipLaunchKernelGGL(k_map_emit, dim3(blocks_for(n_e, kMapThreads)), dim3(kMapThreads), 0, s, n_e, n_f,
                       reinterpret_cast<const int4*>(c->vidx), off_a.p, off_v.p, key_a.p, val_a.p, key_v.p, val_v.p);
    const int bits_a = bits_for(static_cast<unsigned long long>(n_f) * static_cast<unsigned long long>(n_f));
    const int bits_v = bits_for(static_cast<unsigned long long>(n_f));
    size_t tb_a = 0, tb_v = 0;
    if (check_hip(c, hipcub::DeviceRadixSort::SortPairs(nullptr, tb_a, key_a.p, key_a2.p, val_a.p, c->a_cidx,
                                                         static_cast<int>(na), 0, bits_a, s), "sort size") ||
        check_hip(c, hipcub::DeviceRadixSort::SortPairs(nullptr, tb_v, key_v.p, key_v2.p, val_v.p, vlist.p,
                                                         static_cast<int>(nv), 0, bits_v, s), "sort size"))
        return WRAPD_ERR_HIP;
    DevBuf<unsigned char> tmp_sort;
    if (tmp_sort.alloc(c, std::max(tb_a, tb_v))) return WRAPD_ERR_HIP;
    size_t tb = std::max(tb_a, tb_v);
    // A: sorted keys into key_a2, the contributions into a_cidx
    if (check_hip(c, hipcub::DeviceRadixSort::SortPairs(tmp_sort.p, tb, key_a.p, key_a2.p, val_a.p, c->a_cidx,
                                                         static_cast<int>(na), 0, bits_a, s), "sort A"))
        return WRAPD_ERR_HIP;
    // the vertex lists: sorted free vertices into key_v2, the corners into vlist
    tb = std::max(tb_a, tb_v);
    if (check_hip(c, hipcub::DeviceRadixSort::SortPairs(tmp_sort.p, tb, key_v.p, key_v2.p, val_v.p, vlist.p,
                                                         static_cast<int>(nv), 0, bits_v, s), "sort V"))
        return WRAPD_ERR_HIP;
    // runs of A's keys -> slots
    if (head.alloc(c, na) || slot_incl.alloc(c, na)) return WRAPD_ERR_HIP;
    const int gb_a = blocks_for(na, kMapThreads);
    hipLaunchKernelGGL(k_map_heads, dim3(gb_a), dim3(kMapThreads), 0, s, na, key_a2.p, head.p);
    size_t tb_s2 = 0;
    if (check_hip(c, hipcub::DeviceScan::InclusiveSum(nullptr, tb_s2, head.p, slot_incl.p, static_cast<int>(na), s),
                  "scan size"))
        return WRAPD_ERR_HIP;
    DevBuf<unsigned char> tmp_s2;
    if (tmp_s2.alloc(c, tb_s2) ||
        check_hip(c, hipcub::DeviceScan::InclusiveSum(tmp_s2.p, tb_s2, head.p, slot_incl.p, static_cast<int>(na), s),
                  "scan"))
        return WRAPD_ERR_HIP;
    int nnz = 0;
    if (na > 0 && (check_hip(c, hipMemcpyAsync(&nnz, slot_incl.p + (na - 1), sizeof(int), hipMemcpyDeviceToHost, s),
                             "D2H") ||
                   check_hip(c, hipStreamSynchronize(s), "sync")))
        return WRAPD_ERR_HIP;
    c->nnzA = nnz;
    DevBuf<int> scol, d_ap;
    if (alloc_out(c->a_cptr, nnz + 1ull) || srow.alloc(c, nnz) || scol.alloc(c, nnz) || d_ap.alloc(c, n_f + 1ull))
        return WRAPD_ERR_HIP;
    hipLaunchKernelGGL(k_map_slots, dim3(gb_a), dim3(kMapThreads), 0, s, na, n_f, key_a2.p, head.p, slot_incl.p,
                       c->a_cptr, srow.p, scol.p);
    hipLaunchKernelGGL(k_map_fill, dim3(1), dim3(kMapThreads), 0, s, 1ll, static_cast<int>(na), c->a_cptr + nnz);
    hipLaunchKernelGGL(k_map_rowptr, dim3(blocks_for(nnz + 1ll, kMapThreads)), dim3(kMapThreads), 0, s,
                       static_cast<long long>(nnz), n_f, srow.p, d_ap.p);
    // vertex lists: pointers from the sorted free-vertex keys, cpos the inverse of the list
    hipLaunchKernelGGL(k_map_rowptr, dim3(blocks_for(nv + 1, kMapThreads)), dim3(kMapThreads), 0, s, nv, n_f,
                       reinterpret_cast<const int*>(key_v2.p), c->v2e_ptr);
    hipLaunchKernelGGL(k_map_fill, dim3(blocks_for(4ll * n_ep, kMapThreads)), dim3(kMapThreads), 0, s, 4ll * n_ep, -1,
                       c->cpos);
    hipLaunchKernelGGL(k_map_cpos, dim3(blocks_for(nv, kMapThreads)), dim3(kMapThreads), 0, s, nv, vlist.p, c->cpos);
    if (check_hip(c, hipGetLastError(), "mesh map launch")) return WRAPD_ERR_HIP;
    // the pattern for the host's symbolic analysis
    c->Ap.assign(n_f + 1, 0);
    c->Aj.assign(nnz, 0);
    if (check_hip(c, hipMemcpyAsync(c->Ap.data(), d_ap.p, (n_f + 1ull) * sizeof(int), hipMemcpyDeviceToHost, s), "D2H") ||
        check_hip(c, hipMemcpyAsync(c->Aj.data(), scol.p, static_cast<size_t>(nnz) * sizeof(int), hipMemcpyDeviceToHost, s),
                  "D2H") ||
        check_hip(c, hipStreamSynchronize(s), "sync"))
        return WRAPD_ERR_HIP;
    *n_corners = nv;
    return WRAPD_OK;
}

}  // namespace wd
