// LDS FFT passes with the transform length fixed at compile time, for the
// run-time planned fused pipeline (pf_planes.hip).
//
// A pass runs nfft transforms of length L held in LDS as Stockham stages that
// ping-pong between two buffers, one workgroup barrier per stage.  The radix
// plan is chosen at compile time from radices 2..16 (composite ones run in
// registers as nested small DFTs), fewest stages first, so most 2,3,5,7-
// smooth lengths up to 256 need two stages.  Every index and twiddle step is
// a compile-time constant; each thread takes butterflies u = threadIdx.x +
// k blockDim.x of the pass.
//
// Layouts: ROWS — transform f at a[f * st + e] (element stride 1), the
// threads of a row are consecutive; COLS — transform f at a[f + e * st]
// (element stride st), consecutive threads take consecutive transforms at the
// same element (conflict-free LDS columns).
//
// ct_dispatch<DIR, COLS>(L, ...) switches over the lengths listed below and
// returns false for any other length (the caller then runs the run-time
// planned pass).
#pragma once
#include <type_traits>

#include "pf_group_fft.h"

namespace pf {

struct CtPlan {
    int ns = 0;       // stages (-1: no plan)
    int maxr = 0;     // largest radix
    int r[6] = {0, 0, 0, 0, 0, 0};
};

constexpr int kCtRadices[13] = {16, 15, 14, 12, 10, 9, 8, 7, 6, 5, 4, 3, 2};

// Fewest stages, then the smallest largest radix (balanced stages).
constexpr CtPlan ct_plan(int L) {
    CtPlan best;
    if (L == 1) return best;
    best.ns = -1;
    for (int i = 0; i < 13; ++i) {
        const int R = kCtRadices[i];
        if (L % R != 0) continue;
        const CtPlan sub = ct_plan(L / R);
        if (sub.ns < 0 || sub.ns >= 5) continue;
        const int ns = sub.ns + 1;
        const int maxr = sub.maxr > R ? sub.maxr : R;
        if (best.ns < 0 || ns < best.ns || (ns == best.ns && maxr < best.maxr)) {
            best.ns = ns;
            best.maxr = maxr;
            best.r[0] = R;
            for (int s = 0; s < sub.ns; ++s) best.r[s + 1] = sub.r[s];
        }
    }
    return best;
}

constexpr int ct_span(const CtPlan& p, int s) {
    int ns = 1;
    for (int i = 0; i < s; ++i) ns *= p.r[i];
    return ns;
}

// Butterfly j of stage S of the length-L plan: element e of the transform at
// sp[e * es] (input) and dp[e * es] (output).  tw[m] = exp(-2 pi i m / L).
template <int L, int DIR, int S>
PF_HD void ct_bfly(const float2* sp, float2* dp, int j, int es, const float2* tw) {
    constexpr CtPlan P = ct_plan(L);
    constexpr int R = P.r[S];
    constexpr int Ns = ct_span(P, S);
    constexpr int NB = L / R;
    constexpr int STEP = L / (Ns * R);
    const int k = Ns == 1 ? 0 : j % Ns;
    float2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = sp[(j + r * NB) * es];
    if constexpr (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[r] = cmul(v[r], twid<DIR>(tw, r * k * STEP));
    }
    dft_any<R, DIR>(v);
    const int od = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) dp[(od + r * Ns) * es] = v[r];
}

// Sequential run of the whole plan on the host (CPU checks); returns the
// buffer holding the result.
template <int L, int DIR, int S = 0>
inline float2* ct_host_run(float2* a, float2* b, const float2* tw) {
    constexpr CtPlan P = ct_plan(L);
    if constexpr (S >= P.ns) {
        return a;
    } else {
        for (int j = 0; j < L / P.r[S]; ++j) ct_bfly<L, DIR, S>(a, b, j, 1, tw);
        return ct_host_run<L, DIR, S + 1>(b, a, tw);
    }
}

// First-stage input from a loader instead of the source buffer: element e of
// transform f is ld(f, e) (e.g. a pre- or post-processing step computed while
// the stage reads, saving a pass over LDS).
template <int L, int DIR, int S, class LD>
PF_HD void ct_bfly_ld(const LD& ld, int f, float2* dp, int j, int es, const float2* tw) {
    constexpr CtPlan P = ct_plan(L);
    constexpr int R = P.r[S];
    constexpr int NB = L / R;
    static_assert(S == 0, "loaders feed the first stage");
    float2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = ld(f, j + r * NB);
    dft_any<R, DIR>(v);
#pragma unroll
    for (int r = 0; r < R; ++r) dp[(j * R + r) * es] = v[r];
}

// No loader: every stage reads its source buffer.
struct CtNoLoad {};

#if defined(__HIP__)
template <int L, int DIR, bool COLS, int S, class LD>
__device__ __forceinline__ void ct_stage(const float2* a, float2* b, int nfft, int st,
                                         const float2* tw, const LD& ld) {
    constexpr CtPlan P = ct_plan(L);
    constexpr int NB = L / P.r[S];
    constexpr bool kLd = S == 0 && !std::is_same<LD, CtNoLoad>::value;
    const int total = nfft * NB;
    if constexpr (!COLS) {
        for (int u = threadIdx.x; u < total; u += blockDim.x) {
            const int f = u / NB, j = u - f * NB;
            if constexpr (kLd) ct_bfly_ld<L, DIR, S>(ld, f, b + f * st, j, 1, tw);
            else ct_bfly<L, DIR, S>(a + f * st, b + f * st, j, 1, tw);
        }
    } else {
        // u = j nfft + f, stepped without divisions
        int j = (int)threadIdx.x / nfft;
        int f = (int)threadIdx.x - j * nfft;
        const int dj = (int)blockDim.x / nfft, df = (int)blockDim.x - dj * nfft;
        for (int u = threadIdx.x; u < total; u += blockDim.x) {
            if constexpr (kLd) ct_bfly_ld<L, DIR, S>(ld, f, b + f, j, st, tw);
            else ct_bfly<L, DIR, S>(a + f, b + f, j, st, tw);
            j += dj;
            f += df;
            if (f >= nfft) {
                f -= nfft;
                ++j;
            }
        }
    }
}

template <int L, int DIR, bool COLS, int S = 0, class LD = CtNoLoad>
__device__ __forceinline__ float2* ct_run(float2* a, float2* b, int nfft, int st,
                                          const float2* tw, const LD& ld = LD()) {
    constexpr CtPlan P = ct_plan(L);
    static_assert(P.ns >= 0, "no compile-time plan for this length");
    if constexpr (S >= P.ns) {
        return a;
    } else {
        ct_stage<L, DIR, COLS, S>(a, b, nfft, st, tw, ld);
        __syncthreads();
        return ct_run<L, DIR, COLS, S + 1>(b, a, nfft, st, tw, CtNoLoad());
    }
}
#endif

// Lengths with a compiled pass: rows (half the x length of a plane) and
// columns / z (the y and z lengths).  Every 2,3,5,7-smooth length in
// [8, 128] and [16, 256] respectively.
#define PF_CT_ROW_LENGTHS(X) \
    X(8) X(9) X(10) X(12) X(14) X(15) X(16) X(18) X(20) X(21) X(24) X(25) X(27) X(28) X(30) \
    X(32) X(35) X(36) X(40) X(42) X(45) X(48) X(49) X(50) X(54) X(56) X(60) X(63) X(64) X(70) \
    X(72) X(75) X(80) X(81) X(84) X(90) X(96) X(98) X(100) X(105) X(108) X(112) X(120) X(125) \
    X(126) X(128)
#define PF_CT_COL_LENGTHS(X) \
    X(16) X(18) X(20) X(21) X(24) X(25) X(27) X(28) X(30) X(32) X(35) X(36) X(40) X(42) X(45) \
    X(48) X(49) X(50) X(54) X(56) X(60) X(63) X(64) X(70) X(72) X(75) X(80) X(81) X(84) X(90) \
    X(96) X(98) X(100) X(105) X(108) X(112) X(120) X(125) X(126) X(128) X(135) X(140) X(144) \
    X(147) X(150) X(160) X(162) X(168) X(175) X(180) X(189) X(192) X(196) X(200) X(210) X(216) \
    X(224) X(225) X(240) X(243) X(245) X(250) X(252) X(256)

// The odd row lengths (whole odd x lengths, two packed rows per transform).
#define PF_CT_ODD_LENGTHS(X) \
    X(9) X(15) X(21) X(25) X(27) X(35) X(45) X(49) X(63) X(75) X(81) X(105) X(125) X(135) \
    X(147) X(175) X(189) X(225) X(243) X(245)

#if defined(__HIP__)
// Run a compiled pass for length L if there is one: *out receives the buffer
// holding the result.  Column lengths for COLS, else row lengths (the odd
// ones only with ODDROWS).  With a loader the first stage takes its inputs
// from it (a is then only the second ping-pong buffer).
template <int DIR, bool COLS, bool ODDROWS = false, class LD = CtNoLoad>
__device__ __forceinline__ bool ct_dispatch(int L, float2* a, float2* b, int nfft, int st,
                                            const float2* tw, float2** out,
                                            const LD& ld = LD()) {
#define PF_CT_CASE(n) \
    case n: *out = ct_run<n, DIR, COLS, 0, LD>(a, b, nfft, st, tw, ld); return true;
    if constexpr (COLS) {
        switch (L) {
            PF_CT_COL_LENGTHS(PF_CT_CASE)
            default: return false;
        }
    } else if constexpr (ODDROWS) {
        switch (L) {
            PF_CT_ODD_LENGTHS(PF_CT_CASE)
            default: return false;
        }
    } else {
        switch (L) {
            PF_CT_ROW_LENGTHS(PF_CT_CASE)
            default: return false;
        }
    }
#undef PF_CT_CASE
}
#endif

// ---------------------------------------------------------------------------
// Register FFTs of compile-time length L by T threads with stage-by-stage
// butterfly counts (no need for every radix to divide the elements a thread
// holds).  Stage s of radix R has L / R butterflies; thread t runs
// j = t + T b for b < ceil(L / (R T)) (the last ones partly idle) with its R
// inputs in registers v[b R + r] = x[j + r L / R]; the outputs go to an LDS
// exchange buffer at their natural positions and the next stage reads its
// own inputs back.
//
// Register patterns: a transform enters in the input pattern of its first
// radix and leaves in the output pattern of its last, both of the form
// slot b R + r <-> element j + r L / R.  The inverse runs the stages in
// reverse order (REV), so the forward's output pattern is the inverse's
// input pattern and its output pattern the forward's input pattern: a
// forward transform, a pointwise product and the inverse need no exchange
// between them.
struct FlexPlan {
    int ns = -1;
    int T = 0;
    int E = 0;        // register slots per thread (largest stage)
    int cost = 0;
    int r[4] = {0, 0, 0, 0};
};

constexpr int flex_slots(int L, int R, int T) { return ((L / R + T - 1) / T) * R; }

// Cost of an ordered radix list over T threads: thread-slots of work summed
// over stages plus an exchange charge per stage boundary.
constexpr void flex_consider(FlexPlan& best, const int* r, int ns, int L, int emax, int tmax,
                             bool pow2) {
    for (int T = 1; T <= tmax; ++T) {
        if (pow2 && (T & (T - 1)) != 0) continue;
        int E = 0, work = 0;
        for (int s = 0; s < ns; ++s) {
            const int sl = flex_slots(L, r[s], T);
            E = sl > E ? sl : E;
            work += sl;
        }
        if (E > emax) continue;
        const int cost = T * work + (ns - 1) * L;
        if (best.ns < 0 || cost < best.cost || (cost == best.cost && T > best.T)) {
            best.ns = ns;
            best.T = T;
            best.E = E;
            best.cost = cost;
            for (int s = 0; s < 4; ++s) best.r[s] = s < ns ? r[s] : 0;
        }
    }
}

constexpr void flex_search(FlexPlan& best, int* r, int depth, int rem, int L, int emax,
                           int tmax, bool pow2) {
    if (rem == 1) {
        if (depth > 0) flex_consider(best, r, depth, L, emax, tmax, pow2);
        return;
    }
    if (depth == 3) return;
    for (int i = 0; i < 13; ++i) {
        const int R = kCtRadices[i];
        if (rem % R != 0) continue;
        r[depth] = R;
        flex_search(best, r, depth + 1, rem / R, L, emax, tmax, pow2);
    }
}

// Best ordered plan of at most three stages and the thread count for L
// points with at most emax register slots and tmax threads per transform
// (pow2: thread counts that are powers of two only, so blocks of such
// transforms are whole wavefronts).
constexpr FlexPlan flex_plan(int L, int emax, int tmax, bool pow2) {
    FlexPlan best;
    int r[4] = {0, 0, 0, 0};
    flex_search(best, r, 0, L, L, emax, tmax, pow2);
    return best;
}

// Register slots per thread of the run-time planned z stage's transforms:
// the default plan, and an alternative with at most half as many (more
// threads per column, fewer registers per thread) for lengths that have one.
constexpr int kFlexZSlots = 16;
constexpr int kFlexZSlotsAlt = 8;
constexpr bool flex_zalt_ok(int L) { return flex_plan(L, kFlexZSlotsAlt, 32, true).ns >= 1; }

template <int L, int EMAX, int TMAX, bool REV, bool POW2 = false>
struct Flex {
    static constexpr FlexPlan F = flex_plan(L, EMAX, TMAX, POW2);
    static_assert(F.ns >= 1, "no register plan for this length");
    static constexpr int T = F.T, E = F.E, NS = F.ns;
    static constexpr int radix(int s) { return REV ? F.r[NS - 1 - s] : F.r[s]; }
    static constexpr int span(int s) {
        int ns = 1;
        for (int i = 0; i < s; ++i) ns *= radix(i);
        return ns;
    }
    static constexpr int bpt(int s) { return (L / radix(s) + T - 1) / T; }
    // element held in slot b R + r of the input (first stage) / output (last
    // stage) pattern, or -1 for an idle slot
    static PF_HD int in_pos(int t, int slot) {
        constexpr int R = radix(0), NB = L / R;
        const int b = slot / R, r = slot - b * R, j = t + T * b;
        return (slot < bpt(0) * R && j < NB) ? j + r * NB : -1;
    }
    static PF_HD int out_pos(int t, int slot) {
        constexpr int R = radix(NS - 1), NB = L / R;
        const int b = slot / R, r = slot - b * R, j = t + T * b;
        return (slot < bpt(NS - 1) * R && j < NB) ? j + r * NB : -1;
    }
    static constexpr int in_slots() { return bpt(0) * radix(0); }
    static constexpr int out_slots() { return bpt(NS - 1) * radix(NS - 1); }
};

template <int L, int EMAX, int TMAX, bool REV, int DIR, int S, bool POW2 = false>
struct FlexStage {
    using FX = Flex<L, EMAX, TMAX, REV, POW2>;
    static constexpr int T = FX::T, E = FX::E;
    static constexpr bool kLast = S + 1 >= FX::NS;
    static constexpr int R = FX::radix(S), NB = L / R, Ns = FX::span(S);
    static constexpr int STEP = L / (Ns * R), BPT = FX::bpt(S);

    // this stage's butterflies in registers (twiddles, then the R-point DFT)
    static PF_HD void bfly(float2 (&v)[E], int t, const float2* __restrict__ tw) {
#pragma unroll
        for (int b = 0; b < BPT; ++b) {
            const int j = t + T * b;
            if (BPT * T > NB && j >= NB) continue;
            const int k = Ns == 1 ? 0 : j % Ns;
            float2 u[R];
#pragma unroll
            for (int r = 0; r < R; ++r) u[r] = v[b * R + r];
            if constexpr (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) u[r] = cmul(u[r], twid<DIR>(tw, r * k * STEP));
            }
            dft_any<R, DIR>(u);
#pragma unroll
            for (int r = 0; r < R; ++r) v[b * R + r] = u[r];
        }
    }

    // outputs to the exchange buffer at their natural positions
    static PF_HD void write(const float2 (&v)[E], float2* ex, int es, int t) {
#pragma unroll
        for (int b = 0; b < BPT; ++b) {
            const int j = t + T * b;
            if (BPT * T > NB && j >= NB) continue;
            const int k = Ns == 1 ? 0 : j % Ns;
            const int od = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) ex[(od + r * Ns) * es] = v[b * R + r];
        }
    }

    // the next stage's inputs back from the exchange buffer
    static PF_HD void read_next(float2 (&v)[E], const float2* ex, int es, int t) {
        if constexpr (!kLast) {
            constexpr int R2 = FX::radix(S + 1), NB2 = L / R2, BPT2 = FX::bpt(S + 1);
#pragma unroll
            for (int b = 0; b < BPT2; ++b) {
                const int j = t + T * b;
                if (BPT2 * T > NB2 && j >= NB2) continue;
#pragma unroll
                for (int r = 0; r < R2; ++r) v[b * R2 + r] = ex[(j + r * NB2) * es];
            }
        }
    }

#if defined(__HIP__)
    static __device__ __forceinline__ void run(float2 (&v)[E], float2* ex, int es, int t,
                                               const float2* __restrict__ tw, bool active) {
        if (active) bfly(v, t, tw);
        if constexpr (!kLast) {
            if (active) write(v, ex, es, t);
            __syncthreads();
            if (active) read_next(v, ex, es, t);
            __syncthreads();
            FlexStage<L, EMAX, TMAX, REV, DIR, S + 1, POW2>::run(v, ex, es, t, tw, active);
        }
    }
#endif

    // all T threads of one transform in sequence (host checks)
    static void emulate(float2 (*v)[E], float2* ex, const float2* tw) {
        for (int t = 0; t < T; ++t) bfly(v[t], t, tw);
        if constexpr (!kLast) {
            for (int t = 0; t < T; ++t) write(v[t], ex, 1, t);
            for (int t = 0; t < T; ++t) read_next(v[t], ex, 1, t);
            FlexStage<L, EMAX, TMAX, REV, DIR, S + 1, POW2>::emulate(v, ex, tw);
        }
    }
};

#if defined(__HIP__)
// L-point transform of the T threads' registers (input pattern in, output
// pattern out), exchanging through ex[e * es].  Contains workgroup barriers:
// every thread of the block calls it (active = false for idle transforms).
template <int L, int EMAX, int TMAX, bool REV, int DIR, bool POW2, int E>
__device__ __forceinline__ void flex_fft(float2 (&v)[E], float2* ex, int es, int t,
                                         const float2* __restrict__ tw, bool active) {
    static_assert(E == Flex<L, EMAX, TMAX, REV, POW2>::E, "register array of the plan's size");
    FlexStage<L, EMAX, TMAX, REV, DIR, 0, POW2>::run(v, ex, es, t, tw, active);
}
#endif

}  // namespace pf
