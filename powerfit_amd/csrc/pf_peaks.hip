// Peak positions of the best-LCC map: the step after the search
// (Analyzer._watershed, src/powerfit_em/analyzer.py:80-96).  For each cutoff,
// the reference labels the 6-connected components of corr >= cutoff
// (scipy.ndimage.label with its default cross structure, no wrap) and takes
// the position of each component's maximum (scipy.ndimage.maximum_position).
//
// Device form, per cutoff:
//   k_label_init     L[v] = v where corr[v] >= cutoff, else -1;
//   k_label_merge    union-find over the -x, -y, -z neighbours: roots are
//                    hooked with atomicMin (a label only ever decreases, so
//                    every find and every retry loop ends);
//   k_label_flatten  L[v] = root;
//   k_comp_max       per root, the maximum as an order-preserving key;
//   k_comp_argmax    per root, the lowest flat index holding that maximum;
//   k_collect        one position per root, appended.
// Labels are read with agent-scope atomic loads so that hooks made by other
// workgroups are seen.  Among equal maxima of a component the lowest index is
// taken (the reference's choice there follows an unstable argsort).
#include <stdint.h>

#include <algorithm>

#include "../../include/powerfit_hip.h"
#include "pf_internal.h"

namespace pf {
namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ int load_label(const int* L, int x) {
    return __hip_atomic_load(L + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int find_root(const int* L, int x) {
    // L[x] <= x with equality only at roots: the walk strictly decreases
    for (;;) {
        const int p = load_label(L, x);
        if (p == x) return x;
        x = p;
    }
}

__device__ __forceinline__ void unite(int* L, int a, int b) {
    for (;;) {
        a = find_root(L, a);
        b = find_root(L, b);
        if (a == b) return;
        if (a > b) {
            const int t = a;
            a = b;
            b = t;
        }
        // hook root b under a; if b was hooked meanwhile, retry from its new
        // parent (smaller than b)
        const int old = atomicMin(L + b, a);
        if (old == b) return;
        b = old;
    }
}

__device__ __forceinline__ uint64_t order_key(float v) {
    const uint32_t b = __float_as_uint(v);
    return (uint64_t)((b & 0x80000000u) ? ~b : (b | 0x80000000u));
}

__device__ __forceinline__ uint64_t order_key(double v) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | (1ull << 63));
}

template <typename T>
__global__ void k_label_init(const T* __restrict__ corr, int64_t V, T cut, int* __restrict__ L) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V) L[v] = corr[v] >= cut ? (int)v : -1;
}

__global__ void k_label_merge(int* L, int nz, int ny, int nx) {
    const int64_t V = (int64_t)nz * ny * nx;
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V || load_label(L, (int)v) < 0) return;
    const int x = (int)(v % nx);
    const int y = (int)((v / nx) % ny);
    const int z = (int)(v / ((int64_t)nx * ny));
    // a neighbour is foreground iff its label is >= 0 (labels stay >= 0)
    if (x > 0 && load_label(L, (int)v - 1) >= 0) unite(L, (int)v, (int)v - 1);
    if (y > 0 && load_label(L, (int)(v - nx)) >= 0) unite(L, (int)v, (int)(v - nx));
    if (z > 0 && load_label(L, (int)(v - (int64_t)nx * ny)) >= 0)
        unite(L, (int)v, (int)(v - (int64_t)nx * ny));
}

__global__ void k_label_flatten(int* L, int64_t V) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V || load_label(L, (int)v) < 0) return;
    const int r = find_root(L, (int)v);
    __hip_atomic_store(L + v, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_comp_reset(unsigned long long* __restrict__ cmax, int* __restrict__ cpos,
                             int64_t V) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V) {
        cmax[v] = 0ull;   // below the key of every value, -inf included
        cpos[v] = INT32_MAX;
    }
}

template <typename T>
__global__ void k_comp_max(const T* __restrict__ corr, const int* __restrict__ L, int64_t V,
                           unsigned long long* __restrict__ cmax) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const int r = L[v];
    if (r >= 0) atomicMax(cmax + r, (unsigned long long)order_key(corr[v]));
}

template <typename T>
__global__ void k_comp_argmax(const T* __restrict__ corr, const int* __restrict__ L, int64_t V,
                              const unsigned long long* __restrict__ cmax,
                              int* __restrict__ cpos) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const int r = L[v];
    if (r >= 0 && (unsigned long long)order_key(corr[v]) == cmax[r]) atomicMin(cpos + r, (int)v);
}

__global__ void k_collect(const int* __restrict__ L, const int* __restrict__ cpos, int64_t V,
                          unsigned long long* __restrict__ count, int64_t cap,
                          int64_t* __restrict__ out) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V || L[v] != (int)v) return;
    const unsigned long long i = atomicAdd(count, 1ull);
    if ((int64_t)i < cap) out[i] = cpos[v];
}

template <typename T>
int run_peaks(int device, const T* corr, int nz, int ny, int nx, const T* cuts, int ncut,
              int64_t* positions, int64_t capacity, int64_t* count) {
    const int64_t V = (int64_t)nz * ny * nx;
    if (hipSetDevice(device) != hipSuccess) {
        set_error("pf_peak_positions: hipSetDevice failed");
        return kErrHip;
    }
    hipStream_t s;
    if (hipStreamCreate(&s) != hipSuccess) {
        set_error("pf_peak_positions: stream creation failed");
        return kErrHip;
    }
    T* d_corr = nullptr;
    int *d_L = nullptr, *d_cpos = nullptr;
    unsigned long long *d_cmax = nullptr, *d_count = nullptr;
    int64_t* d_out = nullptr;
    const int64_t cap = capacity > 0 ? capacity : 0;
    int st = kOk;
    auto ok = [&](hipError_t e, const char* what) {
        if (st == kOk && e != hipSuccess) {
            set_error(what);
            st = kErrHip;
        }
    };
    ok(hipMalloc((void**)&d_corr, sizeof(T) * V), "pf_peak_positions: allocation failed");
    ok(hipMalloc((void**)&d_L, sizeof(int) * V), "pf_peak_positions: allocation failed");
    ok(hipMalloc((void**)&d_cpos, sizeof(int) * V), "pf_peak_positions: allocation failed");
    ok(hipMalloc((void**)&d_cmax, sizeof(unsigned long long) * V),
       "pf_peak_positions: allocation failed");
    ok(hipMalloc((void**)&d_count, sizeof(unsigned long long)),
       "pf_peak_positions: allocation failed");
    ok(hipMalloc((void**)&d_out, sizeof(int64_t) * (cap > 0 ? cap : 1)),
       "pf_peak_positions: allocation failed");
    if (st == kOk) {
        ok(hipMemcpyAsync(d_corr, corr, sizeof(T) * V, hipMemcpyHostToDevice, s),
           "pf_peak_positions: upload failed");
        ok(hipMemsetAsync(d_count, 0, sizeof(unsigned long long), s),
           "pf_peak_positions: reset failed");
    }
    const dim3 grid((unsigned)((V + kThreads - 1) / kThreads)), block(kThreads);
    for (int c = 0; c < ncut && st == kOk; ++c) {
        hipLaunchKernelGGL(k_label_init<T>, grid, block, 0, s, d_corr, V, cuts[c], d_L);
        hipLaunchKernelGGL(k_label_merge, grid, block, 0, s, d_L, nz, ny, nx);
        hipLaunchKernelGGL(k_label_flatten, grid, block, 0, s, d_L, V);
        hipLaunchKernelGGL(k_comp_reset, grid, block, 0, s, d_cmax, d_cpos, V);
        hipLaunchKernelGGL(k_comp_max<T>, grid, block, 0, s, d_corr, d_L, V, d_cmax);
        hipLaunchKernelGGL(k_comp_argmax<T>, grid, block, 0, s, d_corr, d_L, V, d_cmax, d_cpos);
        hipLaunchKernelGGL(k_collect, grid, block, 0, s, d_L, d_cpos, V, d_count, cap, d_out);
        ok(hipGetLastError(), "pf_peak_positions: launch failed");
    }
    unsigned long long total = 0;
    if (st == kOk) {
        ok(hipMemcpyAsync(&total, d_count, sizeof(total), hipMemcpyDeviceToHost, s),
           "pf_peak_positions: read-back failed");
        ok(hipStreamSynchronize(s), "pf_peak_positions: kernels failed");
    }
    if (st == kOk) {
        const int64_t n = std::min<int64_t>((int64_t)total, cap);
        if (n > 0)
            ok(hipMemcpy(positions, d_out, sizeof(int64_t) * n, hipMemcpyDeviceToHost),
               "pf_peak_positions: read-back failed");
        *count = (int64_t)total;
    }
    (void)hipStreamSynchronize(s);
    (void)hipFree(d_corr);
    (void)hipFree(d_L);
    (void)hipFree(d_cpos);
    (void)hipFree(d_cmax);
    (void)hipFree(d_count);
    (void)hipFree(d_out);
    (void)hipStreamDestroy(s);
    return st;
}

}  // namespace
}  // namespace pf

extern "C" int pf_peak_positions(int device, const void* corr, int dtype, int nz, int ny, int nx,
                                 const void* cutoffs, int ncut, int64_t* positions,
                                 int64_t capacity, int64_t* count) {
    using namespace pf;
    if (!corr || !count || (ncut > 0 && !cutoffs) || ncut < 0 || capacity < 0 ||
        (capacity > 0 && !positions)) {
        set_error("pf_peak_positions: bad arguments");
        return kErrArg;
    }
    if (nz <= 0 || ny <= 0 || nx <= 0 || (int64_t)nz * ny * nx > INT32_MAX) {
        set_error("pf_peak_positions: grid must have 1 .. 2^31 - 1 voxels");
        return kErrArg;
    }
    if (dtype == 0)
        return run_peaks<float>(device, (const float*)corr, nz, ny, nx, (const float*)cutoffs,
                                ncut, positions, capacity, count);
    if (dtype == 1)
        return run_peaks<double>(device, (const double*)corr, nz, ny, nx,
                                 (const double*)cutoffs, ncut, positions, capacity, count);
    set_error("pf_peak_positions: dtype must be 0 (float32) or 1 (float64)");
    return kErrArg;
}
