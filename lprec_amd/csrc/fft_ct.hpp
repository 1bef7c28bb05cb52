// Workgroup FFTs with every size known at compile time, for kernels that the
// library compiles at run time (hipRTC) for the transform lengths of a handle
// that the radix-16 register kernels do not cover (lengths with odd factors,
// e.g. 2064 = 2^4 3 43 for a centre offset at N0 = 2048).
//
// The algorithm is the self-sorting Stockham FFT of fft_lds.hpp (same
// butterflies, same twiddle products, same odd-radix conjugate-pair sums);
// what changes is that the length M, the lines per block NL, their LDS stride
// LDM, the thread count NT and the radix sequence are template arguments, so
// each stage's butterfly count per thread, line and butterfly indices,
// strides and twiddle steps are constants (no run-time divisions, no
// per-stage guards for full rounds), odd-radix weights come from one LDS
// table built once per block, and the stages run out of place between two
// LDS buffers (one barrier per stage, no outputs held across a barrier).
#pragma once
#include "fft_lds.hpp"

namespace lpct {

template <int... Rs>
struct Radices {};

// Conjugate-pair weights of an odd radix R, forward direction:
// wq[(q - 1) GP + p - 1] = exp(-2 pi i (p q mod R) / R), p, q = 1..(R-1)/2,
// GP = ceil(((R-1)/2) / PP) PP columns (zero past (R-1)/2).
#ifndef LPCT_ODD_PP
#define LPCT_ODD_PP 8
#endif
template <int R>
struct OddTable {
    static constexpr int hR = (R - 1) / 2;
    // output pairs per work item: even (the weights are read in pairs), at
    // least 2 and at most LPCT_ODD_PP
    static constexpr int PP = hR <= 2 ? 2 : (hR <= LPCT_ODD_PP ? ((hR + 1) & ~1) : LPCT_ODD_PP);
    static constexpr int G = (hR + PP - 1) / PP;
    static constexpr int GP = G * PP;
    static constexpr int size = hR * GP;
};

// Fill the weight table of odd radix R (run by the whole block once; twM is
// the length-M forward twiddle table, so W_R^m = twM[m M / R]).
template <int M, int R>
__device__ __forceinline__ void odd_table_fill(float2* wq, const float2* __restrict__ twM, int NT) {
    using T = OddTable<R>;
    for (int e = threadIdx.x; e < T::size; e += NT) {
        const int q = e / T::GP + 1, p = e - (q - 1) * T::GP + 1;
        wq[e] = p <= T::hR ? twM[((p * q) % R) * (M / R)] : make_float2(0.f, 0.f);
    }
}

// One closed-form radix-R stage (R in 2, 3, 4, 5, 8) over NL lines of length
// M at LDS stride LDM, Ns = product of the radices before it, from `in` to
// `out` (two LDS buffers: each butterfly is read, transformed and written in
// one pass, one barrier per stage).
template <int M, int NL, int LDM, int NT, int Ns, int R>
__device__ __forceinline__ void stage(const float2* in, float2* out, const float2* __restrict__ tw, int dir, int tid) {
    constexpr int nbf = M / R, total = NL * nbf, NB = (total + NT - 1) / NT;
    constexpr int tstep = M / (Ns * R);
#pragma unroll
    for (int c = 0; c < NB; ++c) {
        const int t = tid + c * NT;
        if ((c + 1) * NT <= total || t < total) {
            const int b = NL == 1 ? 0 : t / nbf, i = t - b * nbf;
            const int base = b * LDM + i;
            float2 v[R];
#pragma unroll
            for (int q = 0; q < R; ++q) v[q] = in[lidx(base + q * nbf)];
            const int k = Ns == 1 ? 0 : i % Ns;
            if (Ns > 1) {
                // w^q from at most three table reads (w, w^2, w^4) and at most
                // two products (the scheme of fft_stage)
                const int e = k * tstep;
                float2 wq[R];
                wq[1] = twiddle(tw, e, dir);
                if (R > 2) wq[2] = twiddle(tw, 2 * e, dir);
                if (R > 3) wq[3] = cmul(wq[1], wq[2]);
                if (R > 4) wq[4] = (R >= 8) ? twiddle(tw, 4 * e, dir) : cmul(wq[2], wq[2]);
                if (R > 5) wq[5] = cmul(wq[1], wq[4]);
                if (R > 6) wq[6] = cmul(wq[2], wq[4]);
                if (R > 7) wq[7] = cmul(wq[3], wq[4]);
#pragma unroll
                for (int q = 1; q < R; ++q) v[q] = cmul(v[q], wq[q]);
            }
            dft_small<R>(v, dir);
            const int dst = b * LDM + (i - k) * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) out[lidx(dst + q * Ns)] = v[q];
        }
    }
    __syncthreads();
}

// Odd radix R > 5 by conjugate pairs (fft_stage_oddsym_t): a work item is one
// butterfly and a group of PP output pairs, weights from the table of R;
// from `in` to `out`, one barrier.
template <int M, int NL, int LDM, int NT, int Ns, int R>
__device__ __forceinline__ void stage_odd(const float2* in, float2* out, const float2* __restrict__ tw,
                                          const float2* __restrict__ wq, int dir, int tid) {
    using T = OddTable<R>;
    constexpr int PP = T::PP, hR = T::hR, G = T::G, GP = T::GP;
    constexpr int nbf = M / R, items = NL * nbf * G, NI = (items + NT - 1) / NT;
    constexpr int tstep = M / (Ns * R);
    const float bsign = dir < 0 ? 1.f : -1.f;   // the table holds the forward weights
#pragma unroll
    for (int c = 0; c < NI; ++c) {
        const int t = tid + c * NT;
        if ((c + 1) * NT <= items || t < items) {
            const int bi = t / G, g = t - bi * G;
            const int b = NL == 1 ? 0 : bi / nbf, i = bi - b * nbf;
            const int k = Ns == 1 ? 0 : i % Ns;
            const int src = b * LDM + i;
            const int dst = b * LDM + (i - k) * R + k;
            const float2 x0 = in[lidx(src)];
            float2 A[PP], B[PP];
#pragma unroll
            for (int j = 0; j < PP; ++j) { A[j] = x0; B[j] = make_float2(0.f, 0.f); }
            float2 S0 = x0;
            const int e1 = k * tstep;
            float2 w1 = make_float2(1.f, 0.f), w1c = w1, wa = w1, wb = w1;
            if (Ns > 1) {
                w1 = twiddle(tw, e1, dir);
                w1c = cconj(w1);
            }
            const float4* wrow = reinterpret_cast<const float4*>(wq + g * PP);
#pragma unroll 3
            for (int q = 1; q <= hR; ++q) {
                float2 xa = in[lidx(src + q * nbf)], xb = in[lidx(src + (R - q) * nbf)];
                if (Ns > 1) {
                    if (((q - 1) & 7) == 0) {
                        wa = twiddle(tw, (q * e1) % M, dir);
                        wb = twiddle(tw, ((R - q) * e1) % M, dir);
                    }
                    xa = cmul(xa, wa);
                    xb = cmul(xb, wb);
                    wa = cmul(wa, w1);
                    wb = cmul(wb, w1c);
                }
                const float2 sq = cadd(xa, xb), dq = csub(xa, xb);
                S0 = cadd(S0, sq);
                const float4* wr = wrow + (q - 1) * (GP / 2);
#pragma unroll
                for (int jj = 0; jj < PP / 2; ++jj) {
                    const float4 w2 = wr[jj];
                    A[2 * jj].x += sq.x * w2.x; A[2 * jj].y += sq.y * w2.x;
                    B[2 * jj].x += dq.x * w2.y; B[2 * jj].y += dq.y * w2.y;
                    A[2 * jj + 1].x += sq.x * w2.z; A[2 * jj + 1].y += sq.y * w2.z;
                    B[2 * jj + 1].x += dq.x * w2.w; B[2 * jj + 1].y += dq.y * w2.w;
                }
            }
            const int p0 = 1 + g * PP;
            if (g == 0) out[lidx(dst)] = S0;
#pragma unroll
            for (int j = 0; j < PP; ++j) {
                const int p = p0 + j;
                if (p <= hR) {
                    const float2 iB = make_float2(-bsign * B[j].y, bsign * B[j].x);
                    out[lidx(dst + p * Ns)] = cadd(A[j], iB);
                    out[lidx(dst + (R - p) * Ns)] = csub(A[j], iB);
                }
            }
        }
    }
    __syncthreads();
}

// The odd radices of a plan, each with its weight table at a fixed offset of
// one LDS array (tables for radices > 5 only).
template <int... Rs>
struct OddTables;
template <>
struct OddTables<> {
    static constexpr int size = 0;
    template <int M>
    __device__ static void fill(float2*, const float2* __restrict__, int) {}
};
template <int R, int... Rest>
struct OddTables<R, Rest...> {
    static constexpr bool odd = R > 5 && (R & 1);
    static constexpr int own = odd ? OddTable<R>::size : 0;
    static constexpr int size = own + OddTables<Rest...>::size;
    template <int M>
    __device__ static void fill(float2* w, const float2* __restrict__ tw, int NT) {
        if constexpr (odd) odd_table_fill<M, R>(w, tw, NT);
        OddTables<Rest...>::template fill<M>(w + own, tw, NT);
    }
};

// Full FFT of NL lines (dir = -1 forward, +1 inverse, unnormalised) through
// the LDS buffers a (holding the input) and b, alternating per stage;
// returns the buffer that holds the result.  wq: the tables of
// OddTables<Rs...>.  Ends with a barrier.
template <int M, int NL, int LDM, int NT, int Ns, int R, int... Rest>
__device__ __forceinline__ float2* run(float2* a, float2* b, const float2* __restrict__ tw,
                                       const float2* __restrict__ wq, int dir, int tid) {
    if constexpr (R > 5 && (R & 1)) {
        stage_odd<M, NL, LDM, NT, Ns, R>(a, b, tw, wq, dir, tid);
        wq += OddTable<R>::size;
    } else {
        stage<M, NL, LDM, NT, Ns, R>(a, b, tw, dir, tid);
    }
    if constexpr (sizeof...(Rest) > 0) return run<M, NL, LDM, NT, Ns * R, Rest...>(b, a, tw, wq, dir, tid);
    else return b;
}

// tid: threadIdx.x, or an equal value the compiler cannot see as loop-
// invariant (then the per-stage index arithmetic is redone per call instead of
// being hoisted into registers held across a caller's loop)
template <int M, int NL, int LDM, int NT, int... Rs>
__device__ __forceinline__ float2* fft(float2* a, float2* b, const float2* __restrict__ tw,
                                       const float2* __restrict__ wq, int dir, Radices<Rs...>, int tid) {
    return run<M, NL, LDM, NT, 1, Rs...>(a, b, tw, wq, dir, tid);
}

}  // namespace lpct
