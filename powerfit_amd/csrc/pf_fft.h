// Mixed-radix FFT building blocks shared by every FFT kernel.
//
// A length-n transform is run as a sequence of Stockham autosort stages
// (radix 2, 3, 4, 5, 7, 8 in closed form, any other prime by an O(p^2) DFT),
// each stage reading one buffer and writing the other so that every
// butterfly is independent within a stage.  The per-butterfly function is
// __host__ __device__: the kernels call it from LDS, and a host build runs
// the same code sequentially to check indexing on the CPU.
//
// Stage s with radix R and span Ns (product of earlier radices), butterfly
// j in [0, n/R):  k = j mod Ns,
//     v[r] = src[j + r n/R] * w^(r k n/(Ns R)),   v = DFT_R(v),
//     dst[(j - k) R + k + r Ns] = v[r].
// Output is in natural order; w = exp(-2 pi i / n) forward, conjugate inverse.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PF_HD __host__ __device__ __forceinline__

namespace pf {

constexpr int kMaxStages = 16;

struct Plan1D {
    int n;                     // transform length
    int nstages;
    int radix[kMaxStages];
};

// Factor n into stages, largest closed-form radices first: 8, 4, 2, 3, 5,
// then remaining primes (7 in closed form, larger ones generic).  Returns false when n < 1.
inline bool make_plan(int n, Plan1D* p) {
    if (n < 1) return false;
    p->n = n;
    p->nstages = 0;
    int m = n;
    while (m % 8 == 0) { p->radix[p->nstages++] = 8; m /= 8; }
    if (m % 4 == 0) { p->radix[p->nstages++] = 4; m /= 4; }
    if (m % 2 == 0) { p->radix[p->nstages++] = 2; m /= 2; }
    const int small[] = {3, 5};
    for (int r : small)
        while (m % r == 0) { p->radix[p->nstages++] = r; m /= r; }
    for (int q = 7; m > 1; q += 2) {
        while (m % q == 0) {
            if (p->nstages >= kMaxStages) return false;
            p->radix[p->nstages++] = q;
            m /= q;
        }
    }
    return true;
}

PF_HD float2 c2(float x, float y) { float2 r; r.x = x; r.y = y; return r; }
PF_HD float2 cadd(float2 a, float2 b) { return c2(a.x + b.x, a.y + b.y); }
PF_HD float2 csub(float2 a, float2 b) { return c2(a.x - b.x, a.y - b.y); }
PF_HD float2 cmul(float2 a, float2 b) {
    return c2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
PF_HD float2 cconj(float2 a) { return c2(a.x, -a.y); }
PF_HD float2 cscale(float2 a, float s) { return c2(a.x * s, a.y * s); }
// multiply by (S * i), S = +-1
template <int S> PF_HD float2 mul_si(float2 a) {
    return S > 0 ? c2(-a.y, a.x) : c2(a.y, -a.x);
}
// twiddle for direction DIR (-1 forward: table value; +1 inverse: conjugate)
template <int DIR> PF_HD float2 twid(const float2* tw, int idx) {
    float2 w = tw[idx];
    return DIR < 0 ? w : cconj(w);
}

// ---- closed-form DFTs of size R; DIR = -1 forward, +1 inverse ----
template <int DIR> PF_HD void dft2(float2* v) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <int DIR> PF_HD void dft3(float2* v) {
    const float h = 0.86602540378443864676f;  // sin(2 pi / 3)
    float2 t1 = cadd(v[1], v[2]);
    float2 t2 = c2(v[0].x - 0.5f * t1.x, v[0].y - 0.5f * t1.y);
    float2 t3 = mul_si<DIR>(cscale(csub(v[1], v[2]), h));
    v[0] = cadd(v[0], t1);
    v[1] = cadd(t2, t3);
    v[2] = csub(t2, t3);
}

template <int DIR> PF_HD void dft4(float2* v) {
    float2 a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
    float2 c = cadd(v[1], v[3]), d = mul_si<DIR>(csub(v[1], v[3]));
    v[0] = cadd(a, c);
    v[2] = csub(a, c);
    v[1] = cadd(b, d);
    v[3] = csub(b, d);
}

template <int DIR> PF_HD void dft5(float2* v) {
    const float c1 = 0.30901699437494742410f;   // cos(2 pi / 5)
    const float c2_ = -0.80901699437494742410f; // cos(4 pi / 5)
    const float s1 = 0.95105651629515357212f;   // sin(2 pi / 5)
    const float s2 = 0.58778525229247312917f;   // sin(4 pi / 5)
    float2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    float2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    float2 a1 = c2(v[0].x + c1 * t1.x + c2_ * t2.x, v[0].y + c1 * t1.y + c2_ * t2.y);
    float2 a2 = c2(v[0].x + c2_ * t1.x + c1 * t2.x, v[0].y + c2_ * t1.y + c1 * t2.y);
    float2 b1 = mul_si<DIR>(c2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y));
    float2 b2 = mul_si<DIR>(c2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y));
    v[0] = cadd(v[0], cadd(t1, t2));
    v[1] = cadd(a1, b1);
    v[4] = csub(a1, b1);
    v[2] = cadd(a2, b2);
    v[3] = csub(a2, b2);
}

template <int DIR> PF_HD void dft8(float2* v) {
    const float r = 0.70710678118654752440f;
    float2 e[4] = {v[0], v[2], v[4], v[6]};
    float2 o[4] = {v[1], v[3], v[5], v[7]};
    dft4<DIR>(e);
    dft4<DIR>(o);
    // o[k] *= w8^k, w8 = (1 + DIR i) / sqrt(2)
    float2 o1 = o[1];
    o[1] = c2(r * (o1.x - (float)DIR * o1.y), r * (o1.y + (float)DIR * o1.x));
    o[2] = mul_si<DIR>(o[2]);
    float2 o3 = o[3];
    o[3] = c2(r * (-o3.x - (float)DIR * o3.y), r * (-o3.y + (float)DIR * o3.x));
    for (int k = 0; k < 4; ++k) {
        v[k] = cadd(e[k], o[k]);
        v[k + 4] = csub(e[k], o[k]);
    }
}

// 7 points from the pairs t_n = v_n + v_(7-n), s_n = v_n - v_(7-n) (n = 1..3):
// X_k = a_k + DIR i b_k and X_(7-k) = a_k - DIR i b_k with
// a_k = v_0 + sum_n cos(2 pi k n / 7) t_n, b_k = sum_n sin(2 pi k n / 7) s_n.
template <int DIR> PF_HD void dft7(float2* v) {
    const float C1 = 0.62348980185873353053f, C2 = -0.22252093395631440429f;
    const float C3 = -0.90096886790241912624f;
    const float S1 = 0.78183148246802980871f, S2 = 0.97492791218182360702f;
    const float S3 = 0.43388373911755812048f;
    const float2 t1 = cadd(v[1], v[6]), t2 = cadd(v[2], v[5]), t3 = cadd(v[3], v[4]);
    const float2 s1 = csub(v[1], v[6]), s2 = csub(v[2], v[5]), s3 = csub(v[3], v[4]);
    const float2 x0 = v[0];
    const float2 a1 = c2(x0.x + C1 * t1.x + C2 * t2.x + C3 * t3.x,
                         x0.y + C1 * t1.y + C2 * t2.y + C3 * t3.y);
    const float2 a2 = c2(x0.x + C2 * t1.x + C3 * t2.x + C1 * t3.x,
                         x0.y + C2 * t1.y + C3 * t2.y + C1 * t3.y);
    const float2 a3 = c2(x0.x + C3 * t1.x + C1 * t2.x + C2 * t3.x,
                         x0.y + C3 * t1.y + C1 * t2.y + C2 * t3.y);
    const float2 b1 = mul_si<DIR>(c2(S1 * s1.x + S2 * s2.x + S3 * s3.x,
                                     S1 * s1.y + S2 * s2.y + S3 * s3.y));
    const float2 b2 = mul_si<DIR>(c2(S2 * s1.x - S3 * s2.x - S1 * s3.x,
                                     S2 * s1.y - S3 * s2.y - S1 * s3.y));
    const float2 b3 = mul_si<DIR>(c2(S3 * s1.x - S1 * s2.x + S2 * s3.x,
                                     S3 * s1.y - S1 * s2.y + S2 * s3.y));
    v[0] = cadd(x0, cadd(t1, cadd(t2, t3)));
    v[1] = cadd(a1, b1);
    v[6] = csub(a1, b1);
    v[2] = cadd(a2, b2);
    v[5] = csub(a2, b2);
    v[3] = cadd(a3, b3);
    v[4] = csub(a3, b3);
}

template <int R, int DIR> PF_HD void dft_fixed(float2* v) {
    if (R == 2) dft2<DIR>(v);
    else if (R == 3) dft3<DIR>(v);
    else if (R == 4) dft4<DIR>(v);
    else if (R == 5) dft5<DIR>(v);
    else if (R == 7) dft7<DIR>(v);
    else if (R == 8) dft8<DIR>(v);
}

// One Stockham butterfly for a closed-form radix R.  Element e of a
// transform lives at base[e * es].  tw holds exp(-2 pi i m / n), m < n.
template <int R, int DIR>
PF_HD void stage_bfly(const float2* sp, float2* dp, int j, int Ns, int n, int es,
                      const float2* tw) {
    const int nb = n / R;
    const int k = j % Ns;
    const int step = n / (Ns * R);
    float2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = sp[(j + r * nb) * es];
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = cmul(v[r], twid<DIR>(tw, r * k * step));
    dft_fixed<R, DIR>(v);
    const int od = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) dp[(od + r * Ns) * es] = v[r];
}

// Stockham butterfly for any radix R (O(R^2)), used for primes >= 11.
template <int DIR>
PF_HD void stage_bfly_generic(const float2* sp, float2* dp, int R, int j, int Ns, int n,
                              int es, const float2* tw) {
    const int nb = n / R;
    const int k = j % Ns;
    const int step = n / (Ns * R);
    const int od = (j - k) * R + k;
    for (int q = 0; q < R; ++q) {
        float2 acc = c2(0.f, 0.f);
        for (int m = 0; m < R; ++m) {
            float2 x = sp[(j + m * nb) * es];
            if (m > 0) x = cmul(x, twid<DIR>(tw, m * k * step));
            const int e = (m * q) % R;
            acc = cadd(acc, e ? cmul(x, twid<DIR>(tw, e * nb)) : x);
        }
        dp[(od + q * Ns) * es] = acc;
    }
}

template <int DIR>
PF_HD void stage_dispatch(const float2* sp, float2* dp, int R, int j, int Ns, int n, int es,
                          const float2* tw) {
    switch (R) {
        case 2: stage_bfly<2, DIR>(sp, dp, j, Ns, n, es, tw); break;
        case 3: stage_bfly<3, DIR>(sp, dp, j, Ns, n, es, tw); break;
        case 4: stage_bfly<4, DIR>(sp, dp, j, Ns, n, es, tw); break;
        case 5: stage_bfly<5, DIR>(sp, dp, j, Ns, n, es, tw); break;
        case 7: stage_bfly<7, DIR>(sp, dp, j, Ns, n, es, tw); break;
        case 8: stage_bfly<8, DIR>(sp, dp, j, Ns, n, es, tw); break;
        default: stage_bfly_generic<DIR>(sp, dp, R, j, Ns, n, es, tw); break;
    }
}

#if defined(__HIP__)
// LDS FFT driver: run all stages of plan p over `nfft` transforms held in LDS
// (element e of transform f at a[f * fs + e * es]), ping-ponging between a
// and b.  Returns the buffer holding the result.  Ends with a barrier.
template <int DIR>
__device__ float2* lds_fft(float2* a, float2* b, const Plan1D& p, const float2* tw, int nfft,
                           int fs, int es) {
    int Ns = 1;
    for (int s = 0; s < p.nstages; ++s) {
        const int R = p.radix[s];
        const int nb = p.n / R;
        const int total = nfft * nb;
        for (int u = threadIdx.x; u < total; u += blockDim.x) {
            int f, j;
            if (es == 1) {
                f = u / nb;
                j = u - f * nb;
            } else {
                j = u / nfft;
                f = u - j * nfft;
            }
            stage_dispatch<DIR>(a + f * fs, b + f * fs, R, j, Ns, p.n, es, tw);
        }
        __syncthreads();
        float2* t = a;
        a = b;
        b = t;
        Ns *= R;
    }
    return a;
}

#endif

}  // namespace pf
