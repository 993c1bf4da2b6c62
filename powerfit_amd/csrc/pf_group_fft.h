// Compile-time register FFTs executed by a group of T threads.
//
// An N-point transform is held by T threads, E = N / T elements each, in the
// strided layout v[e] = x[t + T e] (t = thread index in the group).  Each
// Stockham stage of radix R (R | E) runs E / R butterflies per thread
// entirely from registers — butterfly j = t + T b reads elements
// j + r N / R = t + T (b + r E / R), i.e. v[b + r E / R] — writes its outputs
// to an LDS exchange buffer at their natural positions, and reads the next
// stage's inputs back in the strided layout.  The result is again strided,
// in natural order.  All loops and radices are compile-time, so everything
// unrolls and register arrays are statically indexed.
//
// Every call contains block-wide barriers: all threads of the block must call
// it the same number of times (pass active = false for idle groups).
#pragma once
#include "pf_fft.h"

namespace pf {

struct GPlan {
    int T = 0, E = 0, ns = 0;
    int r[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

// Radix plan for N points over T threads: greedy largest radix dividing both
// the remaining length and E.  ns = -1 when impossible.
constexpr GPlan gplan_for(int N, int T) {
    GPlan p;
    p.T = T;
    if (T < 1 || N % T != 0) {
        p.ns = -1;
        return p;
    }
    p.E = N / T;
    int rem = N;
    // largest first (fewer exchange stages); composite radices run in
    // registers as dft_ab
    const int pref[12] = {16, 24, 12, 20, 8, 10, 6, 4, 2, 3, 5, 7};
    while (rem > 1) {
        int pick = 0;
        for (int i = 0; i < 12; ++i) {
            const int R = pref[i];
            if (rem % R == 0 && p.E % R == 0) {
                pick = R;
                break;
            }
        }
        if (pick == 0 || p.ns == 8) {
            p.ns = -1;
            return p;
        }
        p.r[p.ns++] = pick;
        rem /= pick;
    }
    return p;
}

// Threads per transform for N: fewest stages with E <= emax, ties -> most threads.
constexpr int pick_T(int N, int emax, int tmax) {
    int best_t = 0, best_ns = 100;
    for (int T = tmax; T >= 1; --T) {
        if (N % T != 0 || N / T > emax) continue;
        GPlan p = gplan_for(N, T);
        if (p.ns < 0) continue;
        if (p.ns < best_ns) {
            best_ns = p.ns;
            best_t = T;
        }
    }
    return best_t;
}

// As pick_T but with at least tmin threads per transform (falls back to
// pick_T(N, emax, tmax) when no such plan exists).
constexpr int pick_T_min(int N, int emax, int tmin, int tmax) {
    int best_t = 0, best_ns = 100;
    for (int T = tmax; T >= tmin && T >= 1; --T) {
        if (N % T != 0 || N / T > emax) continue;
        GPlan p = gplan_for(N, T);
        if (p.ns < 0) continue;
        if (p.ns < best_ns) {
            best_ns = p.ns;
            best_t = T;
        }
    }
    return best_t ? best_t : pick_T(N, emax, tmax);
}

// Threads per column for M columns of N points done by NTHR threads: the T
// (E = N / T <= emax) minimising passes x stages x (E + 4) — per-thread work
// per stage plus a fixed exchange cost — with passes = ceil(M / (NTHR / T)).
constexpr int pick_T_cols(int N, int M, int NTHR, int emax) {
    int best_t = 0, best_cost = 1 << 30;
    for (int T = 1; T <= 64 && T <= NTHR; ++T) {
        if (N % T != 0 || N / T > emax || NTHR % T != 0) continue;
        GPlan p = gplan_for(N, T);
        if (p.ns < 0) continue;
        const int cpi = NTHR / T;
        const int passes = (M + cpi - 1) / cpi;
        const int cost = passes * p.ns * (p.E + 4);
        if (cost < best_cost || (cost == best_cost && T > best_t)) {
            best_cost = cost;
            best_t = T;
        }
    }
    return best_t;
}

// 16-point DFT as 4 x 4 with inter-stage twiddles W16^(n2 k1).
template <int DIR>
PF_HD void dft16(float2* v) {
    const float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f;
    const float rr = 0.70710678118654752440f;
    const float d = (float)DIR;
    float2 a[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        float2 q[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
        dft4<DIR>(q);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) a[n2][k1] = q[k1];
    }
    // twiddles W16^m, m = n2 k1 in {1,2,3,2,4,6,3,6,9}
    const float2 w1 = c2(c1, d * s1), w2 = c2(rr, d * rr), w3 = c2(s1, d * c1);
    const float2 w4 = c2(0.f, d), w6 = c2(-rr, d * rr), w9 = c2(-c1, -d * s1);
    a[1][1] = cmul(a[1][1], w1);
    a[1][2] = cmul(a[1][2], w2);
    a[1][3] = cmul(a[1][3], w3);
    a[2][1] = cmul(a[2][1], w2);
    a[2][2] = cmul(a[2][2], w4);
    a[2][3] = cmul(a[2][3], w6);
    a[3][1] = cmul(a[3][1], w3);
    a[3][2] = cmul(a[3][2], w6);
    a[3][3] = cmul(a[3][3], w9);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        float2 q[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
        dft4<DIR>(q);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = q[k2];
    }
}

template <int R, int DIR>
PF_HD void dft_any(float2* v);

// R = A B point DFT in registers (Cooley-Tukey): A-point DFTs over
// x[B n1 + n2], twiddles W_R^(n2 k1), then B-point DFTs -> X[k1 + A k2].
// Lets a thread holding E elements use a composite radix dividing E, so a
// transform needs fewer LDS exchange stages.
template <int A, int B, int DIR>
PF_HD void dft_ab(float2* v) {
    constexpr int R = A * B;
    float2 y[B][A];
#pragma unroll
    for (int n2 = 0; n2 < B; ++n2) {
        float2 u[A];
#pragma unroll
        for (int n1 = 0; n1 < A; ++n1) u[n1] = v[B * n1 + n2];
        dft_any<A, DIR>(u);
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) {
            const int m = (n2 * k1) % R;
            if (m == 0) {
                y[n2][k1] = u[k1];
            } else {
                const double a = (double)DIR * 2.0 * 3.14159265358979323846 * m / R;
                y[n2][k1] = cmul(u[k1], c2((float)__builtin_cos(a), (float)__builtin_sin(a)));
            }
        }
    }
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) {
        float2 w[B];
#pragma unroll
        for (int n2 = 0; n2 < B; ++n2) w[n2] = y[n2][k1];
        dft_any<B, DIR>(w);
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) v[k1 + A * k2] = w[k2];
    }
}

template <int R, int DIR>
PF_HD void dft_any(float2* v) {
    if constexpr (R == 16) dft16<DIR>(v);
    else if constexpr (R == 6) dft_ab<2, 3, DIR>(v);
    else if constexpr (R == 9) dft_ab<3, 3, DIR>(v);
    else if constexpr (R == 10) dft_ab<2, 5, DIR>(v);
    else if constexpr (R == 12) dft_ab<4, 3, DIR>(v);
    else if constexpr (R == 14) dft_ab<2, 7, DIR>(v);
    else if constexpr (R == 15) dft_ab<3, 5, DIR>(v);
    else if constexpr (R == 20) dft_ab<4, 5, DIR>(v);
    else if constexpr (R == 24) dft_ab<8, 3, DIR>(v);
    else dft_fixed<R, DIR>(v);
}

#if defined(__HIP__)
// Synchronisation between the write and read phases of a stage.  WAVE: all
// threads of every group live in one wavefront, whose LDS operations execute
// in issue order, so a wavefront-scope fence (stops compiler reordering)
// suffices; otherwise a workgroup barrier.
template <bool WAVE>
__device__ __forceinline__ void group_sync() {
    if constexpr (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}
#endif

// Exchange slots an N-point transform occupies when PAD > 0 leaves one unused
// slot after every PAD elements (element p at p + p / PAD).
constexpr int pad_extent(int N, int PAD) { return PAD > 0 ? N + (N - 1) / PAD : N; }

// PAD > 0 (contiguous exchanges, es = 1, e.g. the rows of a plane): element p
// of every exchange but the last stage's output of a TO_EX transform sits at
// p + p / PAD.  A stage's butterflies write their R outputs at stride R per
// thread, so without padding the threads of one row, and the rows of one
// wavefront (row stride = 8 mod 32 complex), fall on a few LDS banks; with
// PAD = R the writes spread over all banks.  Needs pad_extent(N, PAD) slots.
template <int N, int T, int DIR, int S, bool WAVE = false, bool TO_EX = false, int PAD = 0>
struct GroupStage {
    static constexpr GPlan P = gplan_for(N, T);
    static constexpr int E = N / T;
    static constexpr bool kLast = S >= P.ns;
    // the last stage of a TO_EX transform leaves its outputs in natural order
    static constexpr bool kNatural = TO_EX && S + 1 >= P.ns;

    static PF_HD int xpos(int p) {
        if constexpr (PAD > 0) return p + p / PAD;
        else return p;
    }

    static constexpr int span() {
        int ns = 1;
        for (int i = 0; i < S; ++i) ns *= P.r[i];
        return ns;
    }

    // Butterflies of this stage from the strided registers; outputs to ex at
    // their natural positions.
    static PF_HD void write_phase(const float2 (&v)[N / T], float2* ex, int es, int t,
                                  const float2* __restrict__ tw) {
        if constexpr (!kLast) {
            constexpr int R = P.r[S];
            constexpr int Ns = span();
            constexpr int B = E / R;
            static_assert(E % R == 0, "radix must divide elements per thread");
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int j = t + T * b;
                const int k = j % Ns;
                float2 u[R];
#pragma unroll
                for (int r = 0; r < R; ++r) u[r] = v[b + r * B];
                if constexpr (Ns > 1) {
#pragma unroll
                    for (int r = 1; r < R; ++r)
                        u[r] = cmul(u[r], twid<DIR>(tw, r * k * (N / (Ns * R))));
                }
                dft_any<R, DIR>(u);
                const int od = (j - k) * R + k;
#pragma unroll
                for (int r = 0; r < R; ++r)
                    ex[(kNatural ? od + r * Ns : xpos(od + r * Ns)) * es] = u[r];
            }
        }
    }

    // Next stage's inputs (strided layout) back from ex.
    static PF_HD void read_phase(float2 (&v)[N / T], const float2* ex, int es, int t) {
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = ex[xpos(t + T * e) * es];
    }

#if defined(__HIP__)
    static __device__ __forceinline__ void run(float2 (&v)[N / T], float2* ex, int es, int t,
                                               const float2* __restrict__ tw, bool active) {
        if constexpr (!kLast) {
            if (active) write_phase(v, ex, es, t, tw);
            // the last stage's butterflies write the outputs at their natural
            // positions: with TO_EX they stay there (no read-back, no sync)
            if constexpr (TO_EX && S + 1 >= P.ns) return;
            group_sync<WAVE>();
            if (active) read_phase(v, ex, es, t);
            group_sync<WAVE>();
            GroupStage<N, T, DIR, S + 1, WAVE, TO_EX, PAD>::run(v, ex, es, t, tw, active);
        }
    }
#endif

    // Sequential emulation of one group (all T threads) for host checks.
    static void emulate(float2 (*v)[N / T], float2* ex, const float2* tw) {
        if constexpr (!kLast) {
            for (int t = 0; t < T; ++t) write_phase(v[t], ex, 1, t, tw);
            for (int t = 0; t < T; ++t) read_phase(v[t], ex, 1, t);
            GroupStage<N, T, DIR, S + 1, WAVE, TO_EX, PAD>::emulate(v, ex, tw);
        }
    }
};

#if defined(__HIP__)
// N-point FFT of the group's strided registers, exchanging through
// ex[0 .. (N-1) * es] (N elements with element stride es).  WAVE = true
// requires every group's T threads to lie in one wavefront (T divides 64 and
// groups are formed from consecutive lanes).
template <int N, int T, int DIR, bool WAVE = false, int PAD = 0>
__device__ __forceinline__ void group_fft(float2 (&v)[N / T], float2* ex, int es, int t,
                                          const float2* __restrict__ tw, bool active) {
    static_assert(gplan_for(N, T).ns >= 0, "no radix plan for this N and T");
    static_assert(!WAVE || (64 % T == 0), "wave-local groups need T | 64");
    GroupStage<N, T, DIR, 0, WAVE, false, PAD>::run(v, ex, es, t, tw, active);
}

// As group_fft, but the transform is left in ex in natural order (element j
// at ex[j * es]) and v holds no defined result.  The last stage ends without
// a synchronisation: readers of ex from other threads must synchronise first
// (a workgroup barrier, or group_sync<true> when they are in the same wave).
template <int N, int T, int DIR, bool WAVE = false, int PAD = 0>
__device__ __forceinline__ void group_fft_ex(float2 (&v)[N / T], float2* ex, int es, int t,
                                             const float2* __restrict__ tw, bool active) {
    static_assert(gplan_for(N, T).ns >= 1, "no radix plan for this N and T");
    static_assert(!WAVE || (64 % T == 0), "wave-local groups need T | 64");
    GroupStage<N, T, DIR, 0, WAVE, true, PAD>::run(v, ex, es, t, tw, active);
}
#endif

}  // namespace pf
