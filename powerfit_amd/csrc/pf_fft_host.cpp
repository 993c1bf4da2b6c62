// Host build of the FFT stage code in pf_fft.h: runs every Stockham stage
// sequentially on the CPU (each butterfly of a stage is independent, so this
// is exactly what a block computes).  Exposed through a tiny C function so
// CPU tests can compare the device butterflies' arithmetic and indexing with
// numpy.fft without a GPU.
#include <cmath>
#include <vector>

#include "pf_fft.h"

extern "C" int pf_host_fft(int n, int dir, const float* in, float* out) {
    pf::Plan1D p;
    if (!pf::make_plan(n, &p)) return -1;
    std::vector<float2> tw(n), a(n), b(n);
    for (int m = 0; m < n; ++m) {
        const double ang = -2.0 * M_PI * (double)m / (double)n;
        tw[m] = pf::c2((float)std::cos(ang), (float)std::sin(ang));
        a[m] = pf::c2(in[2 * m], in[2 * m + 1]);
    }
    float2* src = a.data();
    float2* dst = b.data();
    int Ns = 1;
    for (int s = 0; s < p.nstages; ++s) {
        const int R = p.radix[s];
        for (int j = 0; j < n / R; ++j) {
            if (dir < 0) pf::stage_dispatch<-1>(src, dst, R, j, Ns, n, 1, tw.data());
            else pf::stage_dispatch<1>(src, dst, R, j, Ns, n, 1, tw.data());
        }
        float2* t = src;
        src = dst;
        dst = t;
        Ns *= R;
    }
    for (int m = 0; m < n; ++m) {
        out[2 * m] = src[m].x;
        out[2 * m + 1] = src[m].y;
    }
    return p.nstages;
}

extern "C" int pf_host_plan(int n, int* radix, int max) {
    pf::Plan1D p;
    if (!pf::make_plan(n, &p)) return -1;
    for (int s = 0; s < p.nstages && s < max; ++s) radix[s] = p.radix[s];
    return p.nstages;
}

// Host emulation of the register group FFT (pf_group_fft.h) for the
// (N, T) pairs the fused kernels use.
#include "pf_group_fft.h"

// ex_out: the exchange buffer after the last stage (the natural-order result
// group_fft_ex leaves there) instead of the registers.
template <int N, int T>
static int group_fft_host(int dir, const float* in, float* out, bool ex_out) {
    constexpr int E = N / T;
    std::vector<float2> tw(N), ex(N);
    for (int m = 0; m < N; ++m) {
        const double ang = -2.0 * M_PI * (double)m / (double)N;
        tw[m] = pf::c2((float)std::cos(ang), (float)std::sin(ang));
    }
    std::vector<float2> regs((size_t)T * E);
    auto* v = reinterpret_cast<float2(*)[E]>(regs.data());
    for (int t = 0; t < T; ++t)
        for (int e = 0; e < E; ++e) v[t][e] = pf::c2(in[2 * (t + T * e)], in[2 * (t + T * e) + 1]);
    if (dir < 0) pf::GroupStage<N, T, -1, 0>::emulate(v, ex.data(), tw.data());
    else pf::GroupStage<N, T, 1, 0>::emulate(v, ex.data(), tw.data());
    for (int t = 0; t < T; ++t)
        for (int e = 0; e < E; ++e) {
            const float2 r = ex_out ? ex[t + T * e] : v[t][e];
            out[2 * (t + T * e)] = r.x;
            out[2 * (t + T * e) + 1] = r.y;
        }
    return pf::GroupStage<N, T, -1, 0>::P.ns;
}

#define PF_GROUP_CASE(N, T) \
    if (n == N && t == T) return group_fft_host<N, T>(dir, in, out, ex_out);

static int group_fft_any(int n, int t, int dir, const float* in, float* out, bool ex_out) {
    PF_GROUP_CASE(16, 1) PF_GROUP_CASE(32, 4) PF_GROUP_CASE(32, 2)
    PF_GROUP_CASE(48, 4) PF_GROUP_CASE(64, 8) PF_GROUP_CASE(64, 4)
    PF_GROUP_CASE(96, 4) PF_GROUP_CASE(96, 8) PF_GROUP_CASE(128, 8)
    PF_GROUP_CASE(128, 16) PF_GROUP_CASE(160, 8) PF_GROUP_CASE(192, 8)
    PF_GROUP_CASE(80, 4) PF_GROUP_CASE(40, 4) PF_GROUP_CASE(56, 4) PF_GROUP_CASE(100, 5)
    PF_GROUP_CASE(24, 2) PF_GROUP_CASE(36, 3) PF_GROUP_CASE(72, 6) PF_GROUP_CASE(80, 8)
    return -1;
}

extern "C" int pf_host_group_fft(int n, int t, int dir, const float* in, float* out) {
    return group_fft_any(n, t, dir, in, out, false);
}

extern "C" int pf_host_group_fft_ex(int n, int t, int dir, const float* in, float* out) {
    return group_fft_any(n, t, dir, in, out, true);
}

extern "C" int pf_host_pick_T(int n, int emax, int tmax) { return pf::pick_T(n, emax, tmax); }

// Row FFT with padded exchanges (PAD > 0): the transform of the group's
// registers (to_ex = 0) or, for the TO_EX form, the natural-order result the
// last stage leaves in the exchange buffer (to_ex = 1).
template <int N, int T, int PAD, bool TO_EX>
static int group_fft_pad_host(int dir, const float* in, float* out) {
    constexpr int E = N / T;
    std::vector<float2> tw(N), ex(pf::pad_extent(N, PAD), pf::c2(0.f, 0.f));
    for (int m = 0; m < N; ++m) {
        const double ang = -2.0 * M_PI * (double)m / (double)N;
        tw[m] = pf::c2((float)std::cos(ang), (float)std::sin(ang));
    }
    std::vector<float2> regs((size_t)T * E);
    auto* v = reinterpret_cast<float2(*)[E]>(regs.data());
    for (int t = 0; t < T; ++t)
        for (int e = 0; e < E; ++e) v[t][e] = pf::c2(in[2 * (t + T * e)], in[2 * (t + T * e) + 1]);
    if (dir < 0) pf::GroupStage<N, T, -1, 0, false, TO_EX, PAD>::emulate(v, ex.data(), tw.data());
    else pf::GroupStage<N, T, 1, 0, false, TO_EX, PAD>::emulate(v, ex.data(), tw.data());
    for (int t = 0; t < T; ++t)
        for (int e = 0; e < E; ++e) {
            const int j = t + T * e;
            const float2 r = TO_EX ? ex[j] : v[t][e];
            out[2 * j] = r.x;
            out[2 * j + 1] = r.y;
        }
    return (int)ex.size();
}

#define PF_PAD_CASE(N, T, PAD)                                                              \
    if (n == N && t == T && pad == PAD)                                                     \
        return to_ex ? group_fft_pad_host<N, T, PAD, true>(dir, in, out)                    \
                     : group_fft_pad_host<N, T, PAD, false>(dir, in, out);

extern "C" int pf_host_group_fft_pad(int n, int t, int pad, int to_ex, int dir, const float* in,
                                     float* out) {
    PF_PAD_CASE(32, 4, 8) PF_PAD_CASE(48, 4, 4) PF_PAD_CASE(64, 8, 8) PF_PAD_CASE(80, 4, 8)
    PF_PAD_CASE(96, 4, 16) PF_PAD_CASE(96, 8, 16)
    return -1;
}


// Host run of the compile-time-length LDS passes (pf_ct_fft.h) for every
// length they are compiled for; returns the number of stages, -1 for a length
// without a compiled pass.
#include "pf_ct_fft.h"

template <int L>
static int ct_fft_host(int dir, const float* in, float* out) {
    std::vector<float2> tw(L), a(L), b(L);
    for (int m = 0; m < L; ++m) {
        const double ang = -2.0 * M_PI * (double)m / (double)L;
        tw[m] = pf::c2((float)std::cos(ang), (float)std::sin(ang));
        a[m] = pf::c2(in[2 * m], in[2 * m + 1]);
    }
    const float2* r = dir < 0 ? pf::ct_host_run<L, -1>(a.data(), b.data(), tw.data())
                              : pf::ct_host_run<L, 1>(a.data(), b.data(), tw.data());
    for (int m = 0; m < L; ++m) {
        out[2 * m] = r[m].x;
        out[2 * m + 1] = r[m].y;
    }
    return pf::ct_plan(L).ns;
}

extern "C" int pf_host_ct_fft(int n, int dir, const float* in, float* out) {
    switch (n) {
#define PF_HOST_CASE(L) case L: return ct_fft_host<L>(dir, in, out);
        PF_CT_COL_LENGTHS(PF_HOST_CASE)
        PF_HOST_CASE(8) PF_HOST_CASE(9) PF_HOST_CASE(10) PF_HOST_CASE(12) PF_HOST_CASE(14)
        PF_HOST_CASE(15)
#undef PF_HOST_CASE
        default: return -1;
    }
}

extern "C" int pf_host_ct_plan(int n, int* radix, int max) {
    const pf::CtPlan p = pf::ct_plan(n);
    for (int s = 0; s < p.ns && s < max; ++s) radix[s] = p.r[s];
    return p.ns;
}

// Host run of the register FFTs with per-stage butterfly counts
// (pf_ct_fft.h Flex, EMAX register slots, up to 32 threads per transform, a
// power of two of them: the z stage's parameters): the T threads' registers are filled in the input
// pattern, every stage is emulated thread by thread, and the outputs are read
// from the output pattern.  info[0..3] = T, E, stages, and 1 when the output
// pattern of the forward plan equals the input pattern of the reversed plan
// (and the reversed plan's output the forward's input) slot for slot.
template <int L, int EMAX, bool REV, int DIR>
static void flex_host_run(const float* in, float* out) {
    using FX = pf::Flex<L, EMAX, 32, REV, true>;
    constexpr int T = FX::T, E = FX::E;
    std::vector<float2> tw(L), ex(L);
    for (int m = 0; m < L; ++m) {
        const double ang = -2.0 * M_PI * (double)m / (double)L;
        tw[m] = pf::c2((float)std::cos(ang), (float)std::sin(ang));
    }
    std::vector<float2> regs((size_t)T * E, pf::c2(0.f, 0.f));
    float2(*v)[E] = reinterpret_cast<float2(*)[E]>(regs.data());
    for (int t = 0; t < T; ++t)
        for (int s = 0; s < FX::in_slots(); ++s) {
            const int p = FX::in_pos(t, s);
            if (p >= 0) v[t][s] = pf::c2(in[2 * p], in[2 * p + 1]);
        }
    pf::FlexStage<L, EMAX, 32, REV, DIR, 0, true>::emulate(v, ex.data(), tw.data());
    for (int t = 0; t < T; ++t)
        for (int s = 0; s < FX::out_slots(); ++s) {
            const int p = FX::out_pos(t, s);
            if (p >= 0) {
                out[2 * p] = v[t][s].x;
                out[2 * p + 1] = v[t][s].y;
            }
        }
}

template <int L, int EMAX>
static int flex_host(int dir, int rev, const float* in, float* out, int* info) {
    using F = pf::Flex<L, EMAX, 32, false, true>;
    using B = pf::Flex<L, EMAX, 32, true, true>;
    if (dir < 0)
        rev ? flex_host_run<L, EMAX, true, -1>(in, out) : flex_host_run<L, EMAX, false, -1>(in, out);
    else
        rev ? flex_host_run<L, EMAX, true, 1>(in, out) : flex_host_run<L, EMAX, false, 1>(in, out);
    int same = F::out_slots() == B::in_slots() && B::out_slots() == F::in_slots();
    for (int t = 0; t < F::T && same; ++t) {
        for (int s = 0; s < F::out_slots(); ++s) same &= F::out_pos(t, s) == B::in_pos(t, s);
        for (int s = 0; s < F::in_slots(); ++s) same &= F::in_pos(t, s) == B::out_pos(t, s);
    }
    info[0] = F::T;
    info[1] = F::E;
    info[2] = F::NS;
    info[3] = same;
    return 0;
}

// slots: the z stage's default plan (pf::kFlexZSlots) or its alternative
// (pf::kFlexZSlotsAlt; -2 for a length without one).
extern "C" int pf_host_flex_fft(int n, int slots, int dir, int rev, const float* in, float* out,
                                int* info) {
    switch (n) {
#define PF_HOST_CASE(L)                                                             \
    case L:                                                                         \
        if (slots == pf::kFlexZSlots)                                               \
            return flex_host<L, pf::kFlexZSlots>(dir, rev, in, out, info);          \
        if constexpr (pf::flex_zalt_ok(L))                                          \
            if (slots == pf::kFlexZSlotsAlt)                                        \
                return flex_host<L, pf::kFlexZSlotsAlt>(dir, rev, in, out, info);   \
        return -2;
        PF_CT_COL_LENGTHS(PF_HOST_CASE)
#undef PF_HOST_CASE
        default: return -1;
    }
}
