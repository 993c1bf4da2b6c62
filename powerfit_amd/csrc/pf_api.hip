// C-ABI of libpowerfit_hip.so (declared in include/powerfit_hip.h).
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <deque>
#include <mutex>
#include <vector>

#include "../../include/powerfit_hip.h"
#include "pf_internal.h"
#include "pf_rotate.h"

using namespace pf;

// Work buffers of pf_scan_noalloc: one array per GPU, held from the library's
// first use of that GPU until the process ends.  Host code never uses the
// name as an address: the device address of each GPU's copy comes from
// hipGetSymbolAddress (scratch_ptr below).
constexpr size_t kScratchBytes = PF_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct pf_ctx {
    int device = 0;
    int nz = 0, ny = 0, nx = 0;
    int64_t V = 0, H = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    Fft3D fft{};
    // resident state
    float* d_template = nullptr;  // f32 prepared template
    float* d_mask = nullptr;      // f32 mask weights
    float2* d_F1 = nullptr;       // F[Tf]
    float2* d_F2 = nullptr;       // F[Tf^2]
    uint8_t* d_lccmask = nullptr;
    float2* d_tm = nullptr;       // packed (template, mask) for the fused path
    uint32_t* d_occ = nullptr;    // brick occupancy of d_tm (8^3 bricks)
    unsigned long long* d_keys = nullptr;  // running (LCC, index) keys
    uint32_t flags = 0;
    bool mask_binary = false;
    int chunk_z = 0, chunk_lcc = 0;   // 0 = automatic
    int zsched = 0;                   // z-FFT schedule of the fused z stage
    int support_zr = 0;               // |v| > support_zr: rotated template and mask are zero
    // captured scan (pf_scan_graph)
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int64_t g_first = -1, g_count = -1;
    int g_batch = -1, g_mode = -1, g_cz = -1, g_cl = -1, g_zs = -1;
    int num_cu = 0;
    double* d_rot = nullptr;
    int64_t nrot = 0;
    double norm = 0.0;
    int rmax = -1;
    bool have_target = false, have_template = false;
    // per-batch work buffers
    int cap_real = 0, cap_cplx = 0;
    float* d_real = nullptr;    // [cap][3][V]  (unfused stages only)
    float2* d_cplx = nullptr;   // [cap][3][H]
    // profiling
    bool profile = false;
    struct Rec {
        int which;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    double acc_ms[PF_K_COUNT] = {};
    int64_t acc_n[PF_K_COUNT] = {};
    // progress marks (pf_ctx_mark / pf_ctx_reached): events in ticket order
    struct Mark {
        uint64_t ticket;
        hipEvent_t ev;
    };
    std::deque<Mark> marks;
    std::vector<hipEvent_t> mark_pool;
    uint64_t next_ticket = 1, reached = 0;
};

namespace {

struct DeviceGuard {
    int prev = -1;
    bool ok = false;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) == hipSuccess && hipSetDevice(dev) == hipSuccess) ok = true;
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

#define PF_DEVICE_GUARD(ctx)                                              \
    DeviceGuard guard_((ctx)->device);                                    \
    if (!guard_.ok) {                                                     \
        set_error("cannot select device %d", (ctx)->device);              \
        return kErrHip;                                                   \
    }

int arg_error(const char* msg) {
    set_error("%s", msg);
    return kErrArg;
}

template <typename T>
int dev_alloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) return kOk;
    hipError_t e = hipMalloc((void**)p, sizeof(T) * count);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        set_error("hipMalloc of %zu bytes failed: %s", sizeof(T) * count, hipGetErrorString(e));
        return kErrMemory;
    }
    return kOk;
}

template <typename T>
void dev_free(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

hipEvent_t get_event(pf_ctx* c) {
    if (!c->pool.empty()) {
        hipEvent_t e = c->pool.back();
        c->pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Run f() (which enqueues launches on c->stream), timing it when profiling.
template <typename F>
int timed(pf_ctx* c, int which, F&& f) {
    if (!c->profile) return f();
    hipEvent_t a = get_event(c), b = get_event(c);
    if (!a || !b) {
        set_error("hipEventCreate failed");
        return kErrHip;
    }
    PF_HIP_CHECK(hipEventRecord(a, c->stream));
    int s = f();
    PF_HIP_CHECK(hipEventRecord(b, c->stream));
    c->recs.push_back({which, a, b});
    return s;
}

int drain_profile(pf_ctx* c) {
    if (c->recs.empty()) return kOk;
    PF_HIP_CHECK(hipStreamSynchronize(c->stream));
    for (auto& r : c->recs) {
        float ms = 0.f;
        PF_HIP_CHECK(hipEventElapsedTime(&ms, r.a, r.b));
        c->acc_ms[r.which] += ms;
        c->acc_n[r.which] += 1;
        c->pool.push_back(r.a);
        c->pool.push_back(r.b);
    }
    c->recs.clear();
    return kOk;
}

// Drop the captured scan: its kernel arguments name buffers and values that
// the caller is about to replace (rotations, template and mask, work buffers).
void drop_graph(pf_ctx* c) {
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->graph_exec = nullptr;
    c->graph = nullptr;
    c->g_first = c->g_count = -1;
}

int ensure_capacity(pf_ctx* c, int batch, bool need_real) {
    if (batch > c->cap_cplx || (need_real && batch > c->cap_real)) drop_graph(c);
    if (batch > c->cap_cplx) {
        PF_HIP_CHECK(hipStreamSynchronize(c->stream));
        dev_free(c->d_cplx);
        c->cap_cplx = 0;
        PF_CHECK_STATUS(dev_alloc(&c->d_cplx, (size_t)3 * batch * c->H));
        c->cap_cplx = batch;
    }
    if (need_real && batch > c->cap_real) {
        PF_HIP_CHECK(hipStreamSynchronize(c->stream));
        dev_free(c->d_real);
        c->cap_real = 0;
        PF_CHECK_STATUS(dev_alloc(&c->d_real, (size_t)3 * batch * c->V));
        c->cap_real = batch;
    }
    return kOk;
}

// The run-time planned fused stages: their grid limits, and a rotation
// radius within every axis's half length (signed coordinates per axis).
bool rt_available(const pf_ctx* c) {
    const int half = std::min(c->nz, std::min(c->ny, c->nx)) / 2;
    return plane_rt_supported(c->nz, c->ny, c->nx) && c->rmax <= half;
}

// Scan request after validation: mode resolved (AUTO), batch clamped to the
// range, and which of the three stages run fused.
struct ScanPlan {
    int batch = 0, mode = 0;
    bool fa = false, fb = false, fc = false;   // fused forward / z / inverse stage
    bool rt = false;                            // the fused stages with run-time plans
    bool need_real = false;                     // real work grids (an unfused stage)
};

int plan_scan(pf_ctx* c, int64_t first, int64_t count, int batch, int mode, ScanPlan* sp) {
    if (!c) return arg_error("ctx is null");
    if (!c->have_target || !c->have_template || c->nrot == 0) {
        set_error("First set the template, mask, and rotations.");
        return kErrState;
    }
    if (first < 0 || count < 0 || first + count > c->nrot)
        return arg_error("rotation range outside the rotation set");
    if (mode < PF_MODE_PLAIN || mode > PF_MODE_FUSED_RT) return arg_error("unknown mode");
    // the split-precision FFT stage: the unfused y and z passes, or the fused
    // z stage at the compiled sizes that have its form
    const bool split = (c->flags & PF_FLAG_SPLIT_DFT) != 0;
    const bool fused_ok = fused_supported(c->nz, c->ny, c->nx) && c->rmax <= c->nx / 2 &&
                          (!split || fused_split_supported(c->nx));
    const bool rt_ok = !split && rt_available(c);
    if (mode == PF_MODE_AUTO)
        mode = fused_ok ? PF_MODE_FUSED : (rt_ok ? PF_MODE_FUSED_RT : PF_MODE_BATCHED);
    if ((mode == PF_MODE_FUSED_RT && !rt_ok) ||
        (mode >= PF_MODE_FUSED && mode != PF_MODE_FUSED_RT && !fused_ok)) {
        set_error(split ? "fused pipeline not available with PF_FLAG_SPLIT_DFT"
                        : "fused pipeline not available for grid (%d, %d, %d)",
                  c->nz, c->ny, c->nx);
        return kErrUnsupported;
    }
    if (mode == PF_MODE_PLAIN) batch = 1;
    if (batch < 1 || batch > 4096) return arg_error("batch must be in [1, 4096]");
    // a shorter range needs no more work buffer than its own length (results
    // do not depend on the batch size)
    if (count > 0 && batch > count) batch = (int)count;
    sp->batch = batch;
    sp->mode = mode;
    // stage A: rotate + forward xy; B: z correlate; C: inverse xy + LCC
    sp->rt = mode == PF_MODE_FUSED_RT;
    sp->fa = mode == PF_MODE_FUSED || mode == PF_MODE_MIX_FWD || sp->rt;
    sp->fb = mode == PF_MODE_FUSED || mode == PF_MODE_MIX_Z || sp->rt;
    sp->fc = mode == PF_MODE_FUSED || mode == PF_MODE_MIX_INV || sp->rt;
    sp->need_real = !(sp->fa && sp->fc);
    return kOk;
}

// Enqueue the launches of rotations [first, first + count) on the ctx stream
// with the given work buffers (cplx: 3 * batch * H complex; real: 3 * batch * V
// floats when an unfused stage runs).  Every region of the work buffers a
// launch reads was written earlier in the same batch, so they need no
// clearing.  With `timing`, launches are bracketed by the ctx's profiling
// events when profiling is on.
int enqueue_scan(pf_ctx* c, int64_t first, int64_t count, const ScanPlan& sp, float2* cplx,
                 float* real, bool timing) {
    const int batch = sp.batch;
    const bool fa = sp.fa, fb = sp.fb, fc = sp.fc;
    // one fused stage: the compiled kernels, or their run-time planned form
    auto fused = [&](const FusedArgs& a, int which) {
        return sp.rt ? launch_planes_rt(c->fft, a, which, c->stream)
                     : launch_fused(a, which, c->stream);
    };
    auto run = [&](int which, auto&& f) -> int { return timing ? timed(c, which, f) : f(); };
    const double invV = 1.0 / (double)c->V;
    const int64_t plane = (int64_t)c->ny * c->fft.nxh;
    FusedArgs fa_args{};
    fa_args.n = c->nx;
    fa_args.tm = c->d_tm;
    fa_args.occ = c->d_occ;
    fa_args.rmax = c->rmax;
    fa_args.zr = c->support_zr;
    // split precision: the z stage's default schedule, with its MFMA first stage
    fa_args.split_z = (c->flags & PF_FLAG_SPLIT_DFT) != 0;
    fa_args.zsched = fa_args.split_z ? 0 : c->zsched;
    fa_args.skip_planes = fa && fb && !sp.rt;
    fa_args.work = cplx;
    fa_args.F1 = c->d_F1;
    fa_args.F2 = c->d_F2;
    fa_args.lcc_mask = c->d_lccmask;
    fa_args.keys = c->d_keys;
    fa_args.norm = c->norm;
    fa_args.invV = invV;
    // enough blocks: z kernel ~ ceil(H / (N * 32)) tiles per chunk, LCC kernel N planes
    {
        const int64_t tiles = (c->H / c->nz + 31) / 32;
        const int64_t want = 4LL * c->num_cu;
        const int bz = (int)std::max<int64_t>(1, std::min<int64_t>(batch, tiles * batch / want));
        const int bl = (int)std::max<int64_t>(1, std::min<int64_t>(batch, (int64_t)c->nz * batch / want));
        fa_args.chunk_z = c->chunk_z > 0 ? c->chunk_z : std::min(bz, 16);
        fa_args.chunk_lcc = c->chunk_lcc > 0 ? c->chunk_lcc : std::min(bl, 255);
    }
    for (int64_t k0 = first; k0 < first + count; k0 += batch) {
        const int nb = (int)std::min<int64_t>(batch, first + count - k0);
        fa_args.rots = c->d_rot + 9 * k0;
        fa_args.nrot = nb;
        fa_args.first = k0;
        if (fa) {
            // W^2 = W shortcut only when the z stage is fused (it fills grid 2)
            fa_args.binary_mask = c->mask_binary && fb;
            PF_CHECK_STATUS(run(PF_K_FWD_PLANE, [&] { return fused(fa_args, 0); }));
        } else {
            PF_CHECK_STATUS(run(PF_K_ROTATE, [&] {
                return launch_rotate_pw(c->d_template, c->d_mask, c->d_rot + 9 * k0, nb, c->nz,
                                        c->ny, c->nx, c->rmax, real, c->stream);
            }));
            PF_CHECK_STATUS(run(PF_K_R2C_X, [&] {
                return launch_r2c_x(c->fft, real, cplx, (int64_t)3 * nb * c->nz * c->ny,
                                    c->stream);
            }));
            PF_CHECK_STATUS(run(PF_K_C2C_FWD, [&] {
                return launch_axis(c->fft, 1, cplx, c->fft.nxh, c->fft.nxh,
                                   (int64_t)3 * nb * c->nz, plane, -1, c->stream);
            }));
        }
        if (fb) {
            fa_args.binary_mask = c->mask_binary;
            PF_CHECK_STATUS(run(PF_K_ZCORR, [&] { return fused(fa_args, 1); }));
        } else {
            PF_CHECK_STATUS(run(PF_K_C2C_FWD, [&] {
                return launch_axis(c->fft, 2, cplx, plane, plane, (int64_t)3 * nb,
                                   plane * c->nz, -1, c->stream);
            }));
            PF_CHECK_STATUS(run(PF_K_CONJ_MUL, [&] {
                return launch_conj_mul_pw(cplx, c->d_F1, c->d_F2, c->H, nb, c->stream);
            }));
            PF_CHECK_STATUS(run(PF_K_C2C_INV, [&] {
                return launch_axis(c->fft, 2, cplx, plane, plane, (int64_t)3 * nb,
                                   plane * c->nz, 1, c->stream);
            }));
        }
        if (fc) {
            PF_CHECK_STATUS(run(PF_K_INV_PLANE, [&] { return fused(fa_args, 2); }));
        } else {
            PF_CHECK_STATUS(run(PF_K_C2C_INV, [&] {
                return launch_axis(c->fft, 1, cplx, c->fft.nxh, c->fft.nxh,
                                   (int64_t)3 * nb * c->nz, plane, 1, c->stream);
            }));
            PF_CHECK_STATUS(run(PF_K_C2R_X, [&] {
                return launch_c2r_x(c->fft, cplx, real, (int64_t)3 * nb * c->nz * c->ny,
                                    1.0f, c->stream);
            }));
            PF_CHECK_STATUS(run(PF_K_LCC, [&] {
                return launch_lcc_take_best(real, c->d_lccmask, c->d_keys, c->V, nb, k0,
                                            c->norm, invV, c->stream);
            }));
        }
    }
    return kOk;
}

// Per-device address of `scratch`, resolved once per device with that device
// current (pf_ctx_create calls resolve_scratch for its device).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch[kMaxDevices] = {};

int resolve_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) {
        set_error("device %d beyond the %d the library supports", device, kMaxDevices);
        return kErrArg;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device]) return kOk;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        set_error("hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
        return kErrHip;
    }
    g_scratch[device] = (unsigned char*)p;
    return kOk;
}

unsigned char* scratch_ptr(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch[device];
}

bool radius_ok(int r, int a, int b, int c) {
    return r >= 0 && r <= std::min(a, std::min(b, c)) - 2;
}

}  // namespace

extern "C" {

int pf_abi_version(void) { return PF_ABI_VERSION; }

int pf_device_count(int* count) {
    if (!count) return arg_error("count is null");
    PF_HIP_CHECK(hipGetDeviceCount(count));
    return kOk;
}

int pf_rotate_grid3d(const double* grid, int gz, int gy, int gx, const double* rotmat,
                     int radius, double* out, int oz, int oy, int ox, int nearest) {
    if (!grid || !rotmat || !out) return arg_error("null pointer");
    if (gz < 1 || gy < 1 || gx < 1 || oz < 1 || oy < 1 || ox < 1)
        return arg_error("invalid grid shape");
    if (!radius_ok(radius, gz, gy, gx) || !radius_ok(radius, oz, oy, ox))
        return arg_error("radius must satisfy 0 <= radius <= min(shape) - 2");
    const size_t ng = (size_t)gz * gy * gx, no = (size_t)oz * oy * ox;
    double *dg = nullptr, *dr = nullptr, *dout = nullptr;
    int s = dev_alloc(&dg, ng);
    if (!s) s = dev_alloc(&dr, 9);
    if (!s) s = dev_alloc(&dout, no);
    if (!s) {
        hipError_t e = hipMemcpy(dg, grid, ng * sizeof(double), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(dr, rotmat, 9 * sizeof(double), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(dout, out, no * sizeof(double), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            set_error("rotate_grid3d upload: %s", hipGetErrorString(e));
            s = kErrHip;
        }
    }
    if (!s) s = launch_rotate_f64(dg, gz, gy, gx, dr, radius, dout, oz, oy, ox, nearest, nullptr);
    if (!s) {
        hipError_t e = hipMemcpy(out, dout, no * sizeof(double), hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            set_error("rotate_grid3d download: %s", hipGetErrorString(e));
            s = kErrHip;
        }
    }
    dev_free(dg);
    dev_free(dr);
    dev_free(dout);
    return s;
}

int pf_conj_multiply_dev(const void* in1, const void* in2, void* out, int64_t n, void* stream) {
    if (n < 0) return arg_error("negative length");
    if (n > 0 && (!in1 || !in2 || !out)) return arg_error("null pointer");
    return launch_conj_multiply((const float2*)in1, (const float2*)in2, (float2*)out, n,
                                (hipStream_t)stream);
}

int pf_calc_lcc_dev(const float* gcc, const float* ave, const float* ave2, const uint8_t* mask,
                    float* lcc, int64_t n, void* stream) {
    if (n < 0) return arg_error("negative length");
    if (n > 0 && (!gcc || !ave || !ave2 || !mask || !lcc)) return arg_error("null pointer");
    return launch_calc_lcc(gcc, ave, ave2, mask, lcc, n, (hipStream_t)stream);
}

int pf_rfftn_dev(const float* in, void* out, int nz, int ny, int nx, int batch, void* stream) {
    if (!in || !out) return arg_error("null pointer");
    if (batch < 1) return arg_error("batch must be >= 1");
    Fft3D f;
    PF_CHECK_STATUS(fft3d_init(&f, nz, ny, nx));
    return fft3d_r2c(f, in, (float2*)out, batch, (hipStream_t)stream);
}

int pf_irfftn_dev(void* in, float* out, int nz, int ny, int nx, int batch, void* stream) {
    if (!in || !out) return arg_error("null pointer");
    if (batch < 1) return arg_error("batch must be >= 1");
    Fft3D f;
    PF_CHECK_STATUS(fft3d_init(&f, nz, ny, nx));
    const float scale = (float)(1.0 / ((double)nz * ny * nx));
    return fft3d_c2r(f, (float2*)in, out, batch, scale, (hipStream_t)stream);
}

int pf_ctx_create(int device, int nz, int ny, int nx, uint32_t flags, pf_ctx** out) {
    if (!out) return arg_error("out is null");
    *out = nullptr;
    if (flags & ~(PF_FLAG_NO_BINARY_MASK | PF_FLAG_SPLIT_DFT)) return arg_error("unknown flags");
    if (nz < 1 || ny < 1 || nx < 1) return arg_error("invalid grid shape");
    int ndev = 0;
    PF_HIP_CHECK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return arg_error("no such device");
    pf_ctx* c = new pf_ctx();
    c->device = device;
    c->flags = flags;
    c->nz = nz;
    c->ny = ny;
    c->nx = nx;
    c->V = (int64_t)nz * ny * nx;
    c->H = (int64_t)nz * ny * (nx / 2 + 1);
    DeviceGuard g(device);
    int s = g.ok ? kOk : kErrHip;
    // the scratch of pf_scan_noalloc on this device (with the device current)
    if (!s) s = resolve_scratch(device);
    if (!s) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cu = prop.multiProcessorCount;
        if (c->num_cu <= 0) c->num_cu = 256;
    }
    if (!s) s = fft3d_init(&c->fft, nz, ny, nx);
    // default z schedule: the size's measured best at batch 256 with automatic
    // chunks, the register z stage at 64^3, 72^3, 96^3, 160^3 and 192^3 (K2 per
    // rotation 1.11 -> 1.03, 1.87 -> 1.74, 3.68 -> 3.50, 21.0 -> 17.7 and
    // 33.0 -> 28.2 us)
    if (nz == ny && ny == nx &&
        (nz == 64 || nz == 72 || nz == 96 || nz == 160 || nz == 192))
        c->zsched = 4;
    if (!s && plane_rt_supported(nz, ny, nx)) s = prepare_planes_rt();
    if (!s && (flags & PF_FLAG_SPLIT_DFT)) s = fft3d_set_split(&c->fft, true);
    if (!s && hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
        set_error("hipStreamCreate failed");
        s = kErrHip;
    }
    c->stream = c->own_stream;
    if (!s) s = dev_alloc(&c->d_template, c->V);
    if (!s) s = dev_alloc(&c->d_mask, c->V);
    if (!s) s = dev_alloc(&c->d_F1, c->H);
    if (!s) s = dev_alloc(&c->d_F2, c->H);
    if (!s) s = dev_alloc(&c->d_lccmask, c->V);
    if (!s) s = dev_alloc(&c->d_tm, c->V);
    if (!s) s = dev_alloc(&c->d_occ, (size_t)occupancy_words(nz, ny, nx));
    if (!s) s = dev_alloc(&c->d_keys, c->V);
    if (s) {
        pf_ctx_destroy(c);
        return s;
    }
    *out = c;
    return pf_reset_result(c);
}

int pf_ctx_destroy(pf_ctx* c) {
    if (!c) return kOk;
    {
        DeviceGuard g(c->device);
        if (c->stream) (void)hipStreamSynchronize(c->stream);
        for (auto& r : c->recs) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        for (auto e : c->pool) (void)hipEventDestroy(e);
        for (auto& m : c->marks) (void)hipEventDestroy(m.ev);
        for (auto e : c->mark_pool) (void)hipEventDestroy(e);
        drop_graph(c);
        dev_free(c->d_template);
        dev_free(c->d_mask);
        dev_free(c->d_F1);
        dev_free(c->d_F2);
        dev_free(c->d_lccmask);
        dev_free(c->d_tm);
        dev_free(c->d_occ);
        dev_free(c->d_keys);
        dev_free(c->d_rot);
        dev_free(c->d_real);
        dev_free(c->d_cplx);
        if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    }
    delete c;
    return kOk;
}

int pf_ctx_set_stream(pf_ctx* c, void* stream) {
    if (!c) return arg_error("ctx is null");
    PF_DEVICE_GUARD(c);
    PF_CHECK_STATUS(drain_profile(c));
    PF_HIP_CHECK(hipStreamSynchronize(c->stream));
    c->stream = stream ? (hipStream_t)stream : c->own_stream;
    return kOk;
}

int pf_set_target(pf_ctx* c, const double* target, const uint8_t* lcc_mask) {
    if (!c || !target || !lcc_mask) return arg_error("null pointer");
    PF_DEVICE_GUARD(c);
    double* dt = nullptr;
    float *t32 = nullptr, *t2 = nullptr;
    int s = dev_alloc(&dt, c->V);
    if (!s) s = dev_alloc(&t32, c->V);
    if (!s) s = dev_alloc(&t2, c->V);
    if (!s) {
        hipError_t e = hipMemcpyAsync(dt, target, sizeof(double) * c->V, hipMemcpyHostToDevice,
                                      c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(c->d_lccmask, lcc_mask, c->V, hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) {
            set_error("set_target upload: %s", hipGetErrorString(e));
            s = kErrHip;
        }
    }
    if (!s) s = launch_convert(dt, t32, t2, c->V, c->stream);
    if (!s) s = fft3d_r2c(c->fft, t32, c->d_F1, 1, c->stream);
    if (!s) s = fft3d_r2c(c->fft, t2, c->d_F2, 1, c->stream);
    if (!s && hipStreamSynchronize(c->stream) != hipSuccess) {
        set_error("set_target: stream sync failed");
        s = kErrHip;
    }
    dev_free(dt);
    dev_free(t32);
    dev_free(t2);
    if (!s) c->have_target = true;
    return s;
}

int pf_set_template_mask(pf_ctx* c, const double* tmpl, const double* mask, double norm_factor,
                         int rmax) {
    if (!c || !tmpl || !mask) return arg_error("null pointer");
    if (!(norm_factor > 0)) return arg_error("norm_factor must be > 0 (zero-filled mask)");
    if (!radius_ok(rmax, c->nz, c->ny, c->nx))
        return arg_error("rmax must satisfy 0 <= rmax <= min(shape) - 2");
    PF_DEVICE_GUARD(c);
    PF_HIP_CHECK(hipStreamSynchronize(c->stream));
    drop_graph(c);   // the captured launches carry norm, rmax and the support radius
    double* dt = nullptr;
    int s = dev_alloc(&dt, c->V);
    if (!s && hipMemcpyAsync(dt, tmpl, sizeof(double) * c->V, hipMemcpyHostToDevice,
                             c->stream) != hipSuccess) {
        set_error("set_template_mask upload failed");
        s = kErrHip;
    }
    if (!s) s = launch_convert(dt, c->d_template, nullptr, c->V, c->stream);
    if (!s && hipMemcpyAsync(dt, mask, sizeof(double) * c->V, hipMemcpyHostToDevice,
                             c->stream) != hipSuccess) {
        set_error("set_template_mask upload failed");
        s = kErrHip;
    }
    if (!s) s = launch_convert(dt, c->d_mask, nullptr, c->V, c->stream);
    if (!s) s = launch_pack_tm(c->d_template, c->d_mask, c->d_tm, c->V, c->stream);
    if (!s) s = launch_brick_occupancy(c->d_tm, c->nz, c->ny, c->nx, c->d_occ, c->stream);
    if (!s && hipStreamSynchronize(c->stream) != hipSuccess) {
        set_error("set_template_mask: stream sync failed");
        s = kErrHip;
    }
    dev_free(dt);
    if (!s) {
        bool binary = true;
        for (int64_t i = 0; i < c->V && binary; ++i) binary = (mask[i] == 0.0 || mask[i] == 1.0);
        // support radius of template and mask about voxel 0 (periodic signed
        // coordinates); a rotated voxel v samples corners within sqrt(3) of
        // R^T v, |R^T v| = |v|, so |v| > R + 2 reads only zeros
        int64_t r2max = 0;
        for (int z = 0; z < c->nz; ++z) {
            const int64_t sz = z <= c->nz / 2 ? z : z - c->nz;
            for (int y = 0; y < c->ny; ++y) {
                const int64_t sy = y <= c->ny / 2 ? y : y - c->ny;
                const int64_t base = ((int64_t)z * c->ny + y) * c->nx;
                for (int x = 0; x < c->nx; ++x) {
                    if (tmpl[base + x] == 0.0 && mask[base + x] == 0.0) continue;
                    const int64_t sx = x <= c->nx / 2 ? x : x - c->nx;
                    r2max = std::max(r2max, sz * sz + sy * sy + sx * sx);
                }
            }
        }
        c->support_zr = (int)std::ceil(std::sqrt((double)r2max)) + 2;
        c->mask_binary = binary && !(c->flags & PF_FLAG_NO_BINARY_MASK);
        c->norm = norm_factor;
        c->rmax = rmax;
        c->have_template = true;
    }
    return s;
}

int pf_set_rotations(pf_ctx* c, const double* rot, int64_t count) {
    if (!c || (!rot && count > 0)) return arg_error("null pointer");
    if (count < 0 || count > 0x7fffffffLL) return arg_error("invalid rotation count");
    PF_DEVICE_GUARD(c);
    PF_HIP_CHECK(hipStreamSynchronize(c->stream));
    drop_graph(c);   // the captured launches point into the old rotation array
    dev_free(c->d_rot);
    c->nrot = 0;
    if (count == 0) return kOk;
    PF_CHECK_STATUS(dev_alloc(&c->d_rot, (size_t)count * 9));
    PF_HIP_CHECK(hipMemcpy(c->d_rot, rot, sizeof(double) * 9 * count, hipMemcpyHostToDevice));
    c->nrot = count;
    return kOk;
}

int pf_reset_result(pf_ctx* c) {
    if (!c) return arg_error("ctx is null");
    PF_DEVICE_GUARD(c);
    return launch_fill_keys(c->d_keys, c->V, lcc_key(0.0f, 0u), c->stream);
}

int pf_scan(pf_ctx* c, int64_t first, int64_t count, int batch, int mode) {
    ScanPlan sp;
    PF_CHECK_STATUS(plan_scan(c, first, count, batch, mode, &sp));
    PF_DEVICE_GUARD(c);
    PF_CHECK_STATUS(ensure_capacity(c, sp.batch, sp.need_real));
    return enqueue_scan(c, first, count, sp, c->d_cplx, c->d_real, true);
}

int pf_scan_noalloc(pf_ctx* c, int64_t first, int64_t count, int batch, int mode) {
    ScanPlan sp;
    PF_CHECK_STATUS(plan_scan(c, first, count, batch, mode, &sp));
    const size_t cplx = (size_t)3 * sp.batch * c->H * sizeof(float2);
    const size_t real = sp.need_real ? (size_t)3 * sp.batch * c->V * sizeof(float) : 0;
    // the real grids start at the first 256-byte boundary after the spectra
    const size_t roff = (cplx + 255) & ~(size_t)255;
    const size_t need = sp.need_real ? roff + real : cplx;
    if (need > kScratchBytes) {
        set_error("pf_scan_noalloc: batch %d of a (%d, %d, %d) grid needs %zu work bytes, "
                  "more than the %zu-byte scratch; use a smaller batch",
                  sp.batch, c->nz, c->ny, c->nx, need, kScratchBytes);
        return kErrArg;
    }
    unsigned char* base = scratch_ptr(c->device);
    if (!base) {
        set_error("pf_scan_noalloc: no scratch address for device %d", c->device);
        return kErrState;
    }
    PF_DEVICE_GUARD(c);
    float* rbuf = sp.need_real ? (float*)(base + roff) : nullptr;
    return enqueue_scan(c, first, count, sp, (float2*)base, rbuf, false);
}

int pf_ctx_query(pf_ctx* c) {
    if (!c) return arg_error("ctx is null");
    PF_DEVICE_GUARD(c);
    const hipError_t e = hipStreamQuery(c->stream);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) {
        (void)hipGetLastError();
        return 0;
    }
    set_error("hipStreamQuery: %s", hipGetErrorString(e));
    return kErrHip;
}

int64_t pf_ctx_bytes(int nz, int ny, int nx, int batch, int mode, uint32_t flags, int64_t nrot) {
    if (nz < 1 || ny < 1 || nx < 1 || batch < 1 || nrot < 0)
        return arg_error("invalid shape, batch or rotation count");
    const int64_t V = (int64_t)nz * ny * nx, H = (int64_t)nz * ny * (nx / 2 + 1);
    if (mode == PF_MODE_PLAIN) batch = 1;
    if (nrot > 0 && batch > nrot) batch = (int)nrot;   // pf_scan never uses more
    // the fused work layout holds only the complex spectra; every other mode,
    // and the split-precision FFT, also needs the real grids
    const bool compiled = (mode == PF_MODE_FUSED || mode == PF_MODE_AUTO) &&
                          fused_supported(nz, ny, nx);
    const bool rt = (mode == PF_MODE_FUSED_RT || mode == PF_MODE_AUTO) &&
                    plane_rt_supported(nz, ny, nx);
    const bool split = (flags & PF_FLAG_SPLIT_DFT) != 0;
    const bool fused = (compiled && (!split || fused_split_supported(nx))) || (rt && !split);
    // template, mask (f32), packed pair, keys, lcc mask, two target spectra,
    // brick occupancy, rotations
    int64_t bytes = V * (4 + 4 + 8 + 8 + 1) + 2 * H * 8 + nrot * 9 * 8;
    bytes += 4 * (int64_t)occupancy_words(nz, ny, nx);
    // work buffers (allocated by the first scan) or, before that, the
    // transient buffers of pf_set_target (f64 upload + two f32 grids) and
    // pf_set_template_mask (f64 upload), whichever is larger
    int64_t work = (int64_t)3 * batch * H * 8;
    if (!fused) work += (int64_t)3 * batch * V * 4;
    return bytes + std::max(work, 16 * V);
}

int pf_ctx_mark(pf_ctx* c, uint64_t* ticket) {
    if (!c || !ticket) return arg_error("null pointer");
    PF_DEVICE_GUARD(c);
    hipEvent_t e = nullptr;
    if (!c->mark_pool.empty()) {
        e = c->mark_pool.back();
        c->mark_pool.pop_back();
    } else {
        PF_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    const hipError_t r = hipEventRecord(e, c->stream);
    if (r != hipSuccess) {
        c->mark_pool.push_back(e);
        set_error("hipEventRecord: %s", hipGetErrorString(r));
        return kErrHip;
    }
    *ticket = c->next_ticket++;
    c->marks.push_back({*ticket, e});
    return kOk;
}

int pf_ctx_reached(pf_ctx* c, uint64_t ticket) {
    if (!c) return arg_error("ctx is null");
    if (ticket >= c->next_ticket) return arg_error("unknown ticket");
    PF_DEVICE_GUARD(c);
    // marks complete in stream order: retire them from the front
    while (!c->marks.empty()) {
        const hipError_t e = hipEventQuery(c->marks.front().ev);
        if (e == hipErrorNotReady) {
            (void)hipGetLastError();
            break;
        }
        if (e != hipSuccess) {
            set_error("hipEventQuery: %s", hipGetErrorString(e));
            return kErrHip;
        }
        c->reached = c->marks.front().ticket;
        c->mark_pool.push_back(c->marks.front().ev);
        c->marks.pop_front();
    }
    return ticket <= c->reached ? 1 : 0;
}

int pf_ctx_set_tuning(pf_ctx* c, int chunk_z, int chunk_lcc) {
    if (!c) return arg_error("ctx is null");
    if (chunk_z < 0 || chunk_lcc < 0 || chunk_lcc > 255 || chunk_z > 4096)
        return arg_error("chunk sizes out of range");
    c->chunk_z = chunk_z;
    c->chunk_lcc = chunk_lcc;
    return kOk;
}

int pf_ctx_set_schedule(pf_ctx* c, int z_schedule) {
    if (!c) return arg_error("ctx is null");
    if (z_schedule < 0 || z_schedule >= PF_Z_SCHEDULES) return arg_error("unknown z schedule");
    c->zsched = z_schedule;
    return kOk;
}

int pf_scan_graph(pf_ctx* c, int64_t first, int64_t count, int batch, int mode) {
    if (!c) return arg_error("ctx is null");
    if (count > 0 && batch > count) batch = (int)count;   // as pf_scan does
    if (c->profile) {
        set_error("pf_scan_graph: disable kernel timing first");
        return kErrState;
    }
    PF_DEVICE_GUARD(c);
    const bool same = c->graph_exec && c->g_first == first && c->g_count == count &&
                      c->g_batch == batch && c->g_mode == mode && c->g_cz == c->chunk_z &&
                      c->g_cl == c->chunk_lcc && c->g_zs == c->zsched;
    if (!same) {
        if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
        if (c->graph) (void)hipGraphDestroy(c->graph);
        c->graph_exec = nullptr;
        c->graph = nullptr;
        // validate and size the work buffers outside the capture: a zero-length scan
        PF_CHECK_STATUS(pf_scan(c, first, 0, batch, mode));
        PF_HIP_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int s = pf_scan(c, first, count, batch, mode);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (s) {
            if (g) (void)hipGraphDestroy(g);
            return s;
        }
        if (e != hipSuccess) {
            set_error("hipStreamEndCapture: %s", hipGetErrorString(e));
            return kErrHip;
        }
        c->graph = g;
        PF_HIP_CHECK(hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0));
        c->g_first = first;
        c->g_count = count;
        c->g_batch = batch;
        c->g_mode = mode;
        c->g_cz = c->chunk_z;
        c->g_cl = c->chunk_lcc;
        c->g_zs = c->zsched;
    }
    PF_HIP_CHECK(hipGraphLaunch(c->graph_exec, c->stream));
    return kOk;
}

int pf_ctx_fused_available(pf_ctx* c) {
    if (!c) return arg_error("ctx is null");
    const bool split = (c->flags & PF_FLAG_SPLIT_DFT) != 0;
    if (fused_supported(c->nz, c->ny, c->nx) && (c->rmax < 0 || c->rmax <= c->nx / 2) &&
        (!split || fused_split_supported(c->nx)))
        return 1;
    if (split) return 0;
    return rt_available(c) ? 2 : 0;
}

int pf_get_result(pf_ctx* c, float* best, int32_t* idx) {
    if (!c || !best || !idx) return arg_error("null pointer");
    PF_DEVICE_GUARD(c);
    std::vector<unsigned long long> keys(c->V);
    PF_HIP_CHECK(hipMemcpyAsync(keys.data(), c->d_keys, sizeof(unsigned long long) * c->V,
                                hipMemcpyDeviceToHost, c->stream));
    PF_HIP_CHECK(hipStreamSynchronize(c->stream));
    for (int64_t i = 0; i < c->V; ++i) lcc_unkey(keys[i], best + i, idx + i);
    return kOk;
}

int pf_sync(pf_ctx* c) {
    if (!c) return arg_error("ctx is null");
    PF_DEVICE_GUARD(c);
    PF_HIP_CHECK(hipStreamSynchronize(c->stream));
    return kOk;
}

int pf_ctx_profile(pf_ctx* c, int enable) {
    if (!c) return arg_error("ctx is null");
    PF_DEVICE_GUARD(c);
    PF_CHECK_STATUS(drain_profile(c));
    c->profile = enable != 0;
    for (int i = 0; i < PF_K_COUNT; ++i) {
        c->acc_ms[i] = 0.0;
        c->acc_n[i] = 0;
    }
    return kOk;
}

int pf_ctx_kernel_time(pf_ctx* c, int which, double* ms, int64_t* launches) {
    if (!c || !ms || !launches) return arg_error("null pointer");
    if (which < 0 || which >= PF_K_COUNT) return arg_error("unknown kernel class");
    PF_DEVICE_GUARD(c);
    PF_CHECK_STATUS(drain_profile(c));
    *ms = c->acc_ms[which];
    *launches = c->acc_n[which];
    return kOk;
}

}  // extern "C"
