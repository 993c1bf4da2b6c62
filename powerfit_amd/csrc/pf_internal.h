// Internal declarations shared by the kernel and C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pf_fft.h"

namespace pf {

// Error reporting: the C-ABI returns negative codes and keeps a thread-local
// message for pf_last_error().
enum Status : int {
    kOk = 0,
    kErrArg = -1,       // invalid argument (shape, size, null pointer, order)
    kErrHip = -2,       // HIP runtime error
    kErrState = -3,     // call out of order (e.g. scan before set_target)
    kErrMemory = -4,    // device allocation failed or request over budget
    kErrUnsupported = -5,
};

void set_error(const char* fmt, ...);

#define PF_HIP_CHECK(expr)                                                      \
    do {                                                                        \
        hipError_t e_ = (expr);                                                 \
        if (e_ != hipSuccess) {                                                 \
            ::pf::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                            __FILE__, __LINE__);                                \
            return ::pf::kErrHip;                                               \
        }                                                                       \
    } while (0)

#define PF_CHECK_STATUS(expr)              \
    do {                                   \
        int s_ = (expr);                   \
        if (s_ != 0) return s_;            \
    } while (0)

// Twiddle table exp(-2 pi i m / n), m < n, for the current device (cached).
int twiddles(int n, const float2** out);

// 3-D real <-> half-complex transforms over a batch of volumes, layout
// [batch][nz][ny][nx] real and [batch][nz][ny][nx/2+1] complex (numpy's
// rfftn layout).  The inverse is unnormalised (multiply by 1/V to get
// numpy's irfftn) and overwrites its complex input.
struct Fft3D {
    int nz, ny, nx, nxh;
    Plan1D px, py, pz;         // px: length nx/2 (even nx) or nx (odd nx)
    const float2 *twx, *twy, *twz, *twfull;  // twfull: length nx (R2C post-processing)
    // split-precision mode: the y and z passes do their first radix-16 stage
    // as a DFT-as-GEMM on MFMA (plans py16, pz16 start with radix 16)
    bool split = false;
    Plan1D py16{}, pz16{};
};
int fft3d_init(Fft3D* f, int nz, int ny, int nx);
// Turn the split-precision y/z passes on (both lengths divisible by 16, at
// most kSplitMaxLen) or off.
int fft3d_set_split(Fft3D* f, bool on);
constexpr int kSplitMaxLen = 320;   // two 16-column LDS tiles of complex64: 80 KB at 320
bool make_plan16(int n, Plan1D* p);
int launch_c2c_axis_split(const Plan1D& p, const float2* tw, float2* data, int64_t stride,
                          int64_t inner, int64_t outer, int64_t outer_stride, int dir,
                          hipStream_t s);
// The y (axis 1) or z (axis 2) C2C pass of f, split-precision when f.split.
int launch_axis(const Fft3D& f, int axis, float2* data, int64_t stride, int64_t inner,
                int64_t outer, int64_t outer_stride, int dir, hipStream_t s);
int fft3d_r2c(const Fft3D& f, const float* in, float2* out, int batch, hipStream_t s);
int fft3d_c2r(const Fft3D& f, float2* in, float* out, int batch, float scale, hipStream_t s);

// Individual launches (also used by the batched search pipeline).
int launch_r2c_x(const Fft3D& f, const float* in, float2* out, int64_t rows, hipStream_t s);
int launch_c2r_x(const Fft3D& f, const float2* in, float* out, int64_t rows, float scale,
                 hipStream_t s);
int launch_c2c_axis(const Plan1D& p, const float2* tw, float2* data, int64_t stride,
                    int64_t inner, int64_t outer, int64_t outer_stride, int dir,
                    hipStream_t s);

// Rotation of template (trilinear) and mask (nearest) for `nrot` rotations
// starting at `rots` (row-major 3x3 doubles each).  Writes P, W, W^2 planes
// of batch b at out + (3 b + {0,1,2}) V.  Voxels outside the sphere are 0.
int launch_rotate_pw(const float* tmpl, const float* mask, const double* rots, int nrot,
                     int nz, int ny, int nx, int radius, float* out, hipStream_t s);
// Drop-in rotate_grid3d on FP64 device buffers: writes only in-sphere voxels.
int launch_rotate_f64(const double* grid, int gz, int gy, int gx, const double* rot,
                      int radius, double* out, int oz, int oy, int ox, int nearest,
                      hipStream_t s);

// X[b][g] = conj(X[b][g]) * F_g with F_0 = F_1 = F1, F_2 = F2 (in place).
int launch_conj_mul_pw(float2* X, const float2* F1, const float2* F2, int64_t H, int nrot,
                       hipStream_t s);
int launch_conj_multiply(const float2* a, const float2* b, float2* out, int64_t n,
                         hipStream_t s);

// Fused calc_lcc + take-best over `nrot` rotations whose unnormalised
// correlations sit at real + (3 b + {0,1,2}) V; merges into keys (lcc_key).
int launch_lcc_take_best(const float* real, const uint8_t* mask, unsigned long long* keys,
                         int64_t V, int nrot, int64_t first, double norm, double invV,
                         hipStream_t s);
int launch_fill_keys(unsigned long long* keys, int64_t n, unsigned long long value,
                     hipStream_t s);
int launch_calc_lcc(const float* gcc, const float* ave, const float* ave2,
                    const uint8_t* mask, float* lcc, int64_t n, hipStream_t s);

// ---- fused search path (pf_fused.hip) ----
// True when a compile-time fused pipeline exists for this grid.
bool fused_supported(int nz, int ny, int nx);
struct FusedArgs {
    int n;                       // cubic grid edge
    const float2* tm;            // packed (template, mask), V entries
    const uint32_t* occ;         // brick occupancy of tm (launch_brick_occupancy)
    const double* rots;          // rotation matrices of this batch (row-major 3x3)
    int nrot;                    // rotations in this batch
    int64_t first;               // global index of the first rotation
    int rmax;
    int zr;                      // support half-width: rotated template and mask are zero
                                 // where |signed z| > zr (and rows where z^2+y^2 > zr^2)
    bool skip_planes;            // forward stage skips zero planes and the z stage reads
                                 // them as zero (both stages fused)
    int zsched;                  // z-stage schedule: bit 0 alternative radix plan, bit 1 tile
    bool binary_mask;            // mask values all in {0, 1}: W^2 = W
    float2* work;                // [nrot][3][H] spectra
    const float2* F1;
    const float2* F2;
    const uint8_t* lcc_mask;
    unsigned long long* keys;
    double norm, invV;
    int chunk_z, chunk_lcc;      // rotations per block in the z and LCC kernels
    bool split_z;                // forward z FFTs' first stage as a split-precision
                                 // DFT-GEMM on MFMA (fused_split_supported sizes)
};
// Compiled sizes whose fused z stage has the split-precision MFMA first stage.
bool fused_split_supported(int n);
// Enqueue the three fused stages; `which` selects the stage (0: rotate +
// forward xy, 1: z correlate, 2: inverse xy + LCC) or -1 for all three.
int launch_fused(const FusedArgs& a, int which, hipStream_t s);
int launch_pack_tm(const float* tmpl, const float* mask, float2* tm, int64_t n, hipStream_t s);
// ---- runtime-planned fused path (pf_planes.hip): any grid within its limits ----
bool plane_rt_supported(int nz, int ny, int nx);
// Let the run-time planned kernels take their LDS (function attributes; done
// once per context creation, outside any stream capture).
int prepare_planes_rt();
// Stage `which` (0, 1, 2 as launch_fused) of the runtime-planned pipeline; uses
// f's plans and twiddles and a's buffers and options (a.n, a.occ and
// a.skip_planes are not used: every plane is transformed).
int launch_planes_rt(const Fft3D& f, const FusedArgs& a, int which, hipStream_t s);
// Its z stage with the z length compiled in (pf_zct.hip): lengths with a
// compiled register transform; zr as for the LDS-planned z stage.
bool zcorr_ct_supported(int nz);
int launch_zcorr_ct(const Fft3D& f, const FusedArgs& a, int zr, hipStream_t s);

// 1 bit per 8^3 brick of an interleaved grid (ceil(n / 8) bricks per axis):
// occupancy_words(nz, ny, nx) words.
int occupancy_words(int nz, int ny, int nx);
int launch_brick_occupancy(const float2* tm, int nz, int ny, int nx, uint32_t* occ,
                           hipStream_t s);

// f64 -> f32 conversion, optionally also writing the square (computed in f64).
int launch_convert(const double* in, float* out, float* out_sq, int64_t n, hipStream_t s);

}  // namespace pf
