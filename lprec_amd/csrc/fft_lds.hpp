// Workgroup-cooperative FP32 FFTs held in LDS (gfx950, wave64).
//
// A line of M complex values lives at buf[lidx(line * ldm + i)].  Each stage is a
// self-sorting Stockham radix-R pass done in place: every thread first pulls
// the inputs of its butterflies into registers, the workgroup synchronises,
// then the butterflies are evaluated and scattered to their output slots.
// Radices 2, 3, 4, 5 and 8 have closed-form butterflies; any other factor
// (7, 11, 43, ...) runs through the generic O(R) per-output stage.  Twiddles
// come from a table tw[j] = exp(-2 pi i j / M) built on the host in FP64.
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

#define LP_MAX_STAGES 16

struct FftPlan {
    int M;
    int nst;
    int radix[LP_MAX_STAGES];
};

// LDS slot of complex element x of an FFT buffer: one pad slot after every 16
// elements, so the Stockham scatter (stride R) and the transposed row/column
// exchanges spread over all banks.
__device__ __forceinline__ int lidx(int x) { return x + (x >> 4); }
// complex slots needed for n elements
__host__ __device__ constexpr int lds_slots(int n) { return n + (n >> 4) + 1; }

// Quotient of 0 <= t < 2^22 by 0 < d < 2^24 without an integer division:
// (t + 0.5) / d lies at least 0.5 / d from an integer, and the float product
// (t + 0.5) * fl(1 / d) errs by at most ((t + 0.5) / d) 2^-23, less than that
// distance while t + 0.5 < 2^22.
struct FDiv {
    int d;
    float inv;
    __device__ __forceinline__ explicit FDiv(int d_) : d(d_), inv(1.0f / (float)d_) {}
    __device__ __forceinline__ int quo(int t) const { return (int)(((float)t + 0.5f) * inv); }
    __device__ __forceinline__ int rem(int t) const { return t - quo(t) * d; }
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// multiply by -i (dir = -1) or +i (dir = +1)
__device__ __forceinline__ float2 rot90(float2 a, int dir) {
    return dir < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

template <int R>
__device__ __forceinline__ void dft_small(float2* v, int dir);

template <>
__device__ __forceinline__ void dft_small<2>(float2* v, int) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <>
__device__ __forceinline__ void dft_small<4>(float2* v, int dir) {
    float2 a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
    float2 c = cadd(v[1], v[3]), d = rot90(csub(v[1], v[3]), dir);
    v[0] = cadd(a, c);
    v[2] = csub(a, c);
    v[1] = cadd(b, d);
    v[3] = csub(b, d);
}

template <>
__device__ __forceinline__ void dft_small<8>(float2* v, int dir) {
    float2 e[4] = {v[0], v[2], v[4], v[6]};
    float2 o[4] = {v[1], v[3], v[5], v[7]};
    dft_small<4>(e, dir);
    dft_small<4>(o, dir);
    const float r = 0.70710678118654752440f;
    // o[k] *= W8^k, W8 = exp(dir * 2 pi i / 8)
    float2 w1 = make_float2(r, dir * r);
    o[1] = cmul(o[1], w1);
    o[2] = rot90(o[2], dir);
    o[3] = cmul(o[3], make_float2(-r, dir * r));
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] = cadd(e[k], o[k]);
        v[k + 4] = csub(e[k], o[k]);
    }
}

template <>
__device__ __forceinline__ void dft_small<3>(float2* v, int dir) {
    const float c = -0.5f, s = 0.86602540378443864676f * dir;
    float2 a = v[0], b = v[1], d = v[2];
    float2 t = cadd(b, d);
    float2 u = csub(b, d);
    float2 m = make_float2(a.x + c * t.x, a.y + c * t.y);
    float2 su = make_float2(-s * u.y, s * u.x);  // i * s * u
    v[0] = cadd(a, t);
    v[1] = cadd(m, su);
    v[2] = csub(m, su);
}

template <>
__device__ __forceinline__ void dft_small<5>(float2* v, int dir) {
    const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
    const float s1 = 0.95105651629515357212f * dir, s2 = 0.58778525229247312917f * dir;
    float2 a = v[0];
    float2 t1 = cadd(v[1], v[4]), u1 = csub(v[1], v[4]);
    float2 t2 = cadd(v[2], v[3]), u2 = csub(v[2], v[3]);
    float2 m1 = make_float2(a.x + c1 * t1.x + c2 * t2.x, a.y + c1 * t1.y + c2 * t2.y);
    float2 m2 = make_float2(a.x + c2 * t1.x + c1 * t2.x, a.y + c2 * t1.y + c1 * t2.y);
    // i*(s1 u1 + s2 u2) and i*(s2 u1 - s1 u2)
    float2 n1 = make_float2(s1 * u1.x + s2 * u2.x, s1 * u1.y + s2 * u2.y);
    float2 n2 = make_float2(s2 * u1.x - s1 * u2.x, s2 * u1.y - s1 * u2.y);
    float2 in1 = make_float2(-n1.y, n1.x), in2 = make_float2(-n2.y, n2.x);
    v[0] = make_float2(a.x + t1.x + t2.x, a.y + t1.y + t2.y);
    v[1] = cadd(m1, in1);
    v[4] = csub(m1, in1);
    v[2] = cadd(m2, in2);
    v[3] = csub(m2, in2);
}

__device__ __forceinline__ float2 twiddle(const float2* __restrict__ tw, int e, int dir) {
    float2 w = tw[e];
    return dir < 0 ? w : make_float2(w.x, -w.y);
}

// One closed-form radix-R Stockham stage over `nlines` lines.  Requires
// nlines * M <= EPT * blockDim.x.
template <int R, int EPT>
__device__ void fft_stage(float2* buf, int ldm, int nlines, int M, int Ns, const float2* __restrict__ tw, int dir) {
    constexpr int MAXB = (EPT + R - 1) / R;
    const int nbf = M / R;
    const int total = nlines * nbf;
    const int nt = blockDim.x;
    const FDiv dnbf(nbf), dNs(Ns);
    float2 v[MAXB][R];
#pragma unroll
    for (int c = 0; c < MAXB; ++c) {
        int t = threadIdx.x + c * nt;
        if (t < total) {
            int b = dnbf.quo(t), i = t - b * nbf;
            const int base = b * ldm + i;
#pragma unroll
            for (int q = 0; q < R; ++q) v[c][q] = buf[lidx(base + q * nbf)];
        }
    }
    __syncthreads();
    const int tstep = M / (Ns * R);
#pragma unroll
    for (int c = 0; c < MAXB; ++c) {
        int t = threadIdx.x + c * nt;
        if (t < total) {
            int b = dnbf.quo(t), i = t - b * nbf;
            int k = dNs.rem(i);
            if (Ns > 1) {
                // w^q from at most three table reads (w, w^2, w^4) and at most
                // two products, so every twiddle stays within a few ulp
                int base = k * tstep;
                float2 w1 = twiddle(tw, base, dir);
                float2 wq[R];
                wq[1] = w1;
                if (R > 2) wq[2] = twiddle(tw, 2 * base, dir);
                if (R > 3) wq[3] = cmul(w1, wq[2]);
                if (R > 4) wq[4] = (R >= 8) ? twiddle(tw, 4 * base, dir) : cmul(wq[2], wq[2]);
                if (R > 5) wq[5] = cmul(w1, wq[4]);
                if (R > 6) wq[6] = cmul(wq[2], wq[4]);
                if (R > 7) wq[7] = cmul(wq[3], wq[4]);
#pragma unroll
                for (int q = 1; q < R; ++q) v[c][q] = cmul(v[c][q], wq[q]);
            }
            dft_small<R>(v[c], dir);
            const int dst = b * ldm + (i - k) * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) buf[lidx(dst + q * Ns)] = v[c][q];
        }
    }
    __syncthreads();
}

// Generic radix-R stage: each work item produces one output of one butterfly
// by an R-term sum.  Requires nlines * M <= EPT * blockDim.x.
template <int EPT>
__device__ void fft_stage_generic(float2* buf, int ldm, int nlines, int M, int Ns, int R,
                                  const float2* __restrict__ tw, int dir) {
    const int nbf = M / R;
    const int total = nlines * M;
    const int nt = blockDim.x;
    const int tstep = M / (Ns * R);
    const FDiv dM(M), dnbf(nbf), dNs(Ns);
    float2 out[EPT];
#pragma unroll
    for (int c = 0; c < EPT; ++c) {
        int t = threadIdx.x + c * nt;
        if (t < total) {
            int b = dM.quo(t), o = t - b * M;
            int p = dnbf.quo(o), i = o - p * nbf;
            int k = dNs.rem(i);
            int e = tstep * (k + p * Ns);  // exponent of W_M per input index q
            const int src = b * ldm + i;
            float2 acc = make_float2(0.f, 0.f);
            // W^(e q) by recurrence, re-read from the table every 8 terms
            // (at most 7 chained products: ~4e-7 relative)
            const float2 w1 = twiddle(tw, e, dir);
            float2 w = make_float2(1.f, 0.f);
            int ex = 0;
            for (int q = 0; q < R; ++q) {
                if ((q & 7) == 0) w = twiddle(tw, ex, dir);
                float2 x = buf[lidx(src + q * nbf)];
                acc.x += x.x * w.x - x.y * w.y;
                acc.y += x.x * w.y + x.y * w.x;
                w = cmul(w, w1);
                ex += e;
                if (ex >= M) ex -= M;
            }
            out[c] = acc;
        }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < EPT; ++c) {
        int t = threadIdx.x + c * nt;
        if (t < total) {
            int b = dM.quo(t), o = t - b * M;
            int p = dnbf.quo(o), i = o - p * nbf;
            int k = dNs.rem(i);
            buf[lidx(b * ldm + (i - k) * R + k + p * Ns)] = out[c];
        }
    }
    __syncthreads();
}

// Odd radix R > 5 by conjugate pairs: with x'_q the twiddled inputs,
// s_q = x'_q + x'_{R-q}, d_q = x'_q - x'_{R-q} (q = 1..(R-1)/2) and
// w_m = exp(dir 2 pi i m / R),
//   y_0 = x'_0 + sum s_q,  y_p = A_p + i B_p,  y_{R-p} = A_p - i B_p,
//   A_p = x'_0 + sum_q s_q Re w_{pq},  B_p = sum_q d_q Im w_{pq}.
// A work item is one butterfly and a group of up to PP output pairs, so each
// input is read and twiddled once per group instead of once per output.
// Returns false (doing nothing) when a thread would need more than MAXW work
// items; the caller then uses fft_stage_generic.
template <int EPT>
__device__ bool fft_stage_oddsym(float2* buf, int ldm, int nlines, int M, int Ns, int R,
                                 const float2* __restrict__ tw, int dir) {
    constexpr int PP = 8, MAXW = 1, MAXR = 128;
    const int nbf = M / R, hR = (R - 1) / 2;
    const int G = (hR + PP - 1) / PP;
    const int total = nlines * nbf * G;
    const int nt = blockDim.x;
    if (R > MAXR || (R & 1) == 0 || total > MAXW * nt) return false;
    __shared__ float2 wtab[MAXR];   // w_m, m < R
    for (int m = threadIdx.x; m < R; m += nt) wtab[m] = twiddle(tw, m * nbf, dir);
    __syncthreads();
    const int tstep = M / (Ns * R);
    float2 ya[MAXW][PP], yb[MAXW][PP], y0[MAXW];
#pragma unroll
    for (int c = 0; c < MAXW; ++c) {
        int t = threadIdx.x + c * nt;
        if (t < total) {
            const FDiv dG(G), dnbf(nbf), dNs(Ns);
            int bi = dG.quo(t), g = t - bi * G;
            int b = dnbf.quo(bi), i = bi - b * nbf;
            int k = dNs.rem(i);
            const int src = b * ldm + i;
            const int p0 = 1 + g * PP;
            const float2 x0 = buf[lidx(src)];
            float2 A[PP], B[PP];
            int mq[PP];
#pragma unroll
            for (int j = 0; j < PP; ++j) { A[j] = x0; B[j] = make_float2(0.f, 0.f); mq[j] = 0; }
            float2 S0 = x0;
            // twiddles W^(q e1) (ascending q) and W^((R-q) e1) (descending), by
            // recurrence re-read from the table every 8 terms
            const int e1 = k * tstep;
            const float2 w1 = twiddle(tw, e1, dir);
            const float2 w1c = make_float2(w1.x, -w1.y);
            float2 wa = make_float2(1.f, 0.f), wb = make_float2(1.f, 0.f);
            for (int q = 1; q <= hR; ++q) {
                float2 xa = buf[lidx(src + q * nbf)], xb = buf[lidx(src + (R - q) * nbf)];
                if (Ns > 1) {
                    if (((q - 1) & 7) == 0) {
                        wa = twiddle(tw, (int)(((long long)q * e1) % M), dir);
                        wb = twiddle(tw, (int)(((long long)(R - q) * e1) % M), dir);
                    }
                    xa = cmul(xa, wa);
                    xb = cmul(xb, wb);
                    wa = cmul(wa, w1);
                    wb = cmul(wb, w1c);
                }
                const float2 sq = cadd(xa, xb), dq = csub(xa, xb);
                S0 = cadd(S0, sq);
#pragma unroll
                for (int j = 0; j < PP; ++j) {
                    // m = (p q) mod R, advanced by p each step
                    int m = mq[j] + p0 + j;
                    m = m >= R ? m - R : m;
                    mq[j] = m;
                    const float2 w = wtab[m];
                    A[j].x += sq.x * w.x; A[j].y += sq.y * w.x;
                    B[j].x += dq.x * w.y; B[j].y += dq.y * w.y;
                }
            }
            y0[c] = S0;
#pragma unroll
            for (int j = 0; j < PP; ++j) {
                const float2 iB = make_float2(-B[j].y, B[j].x);
                ya[c][j] = cadd(A[j], iB);
                yb[c][j] = csub(A[j], iB);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < MAXW; ++c) {
        int t = threadIdx.x + c * nt;
        if (t < total) {
            const FDiv dG(G), dnbf(nbf), dNs(Ns);
            int bi = dG.quo(t), g = t - bi * G;
            int b = dnbf.quo(bi), i = bi - b * nbf;
            int k = dNs.rem(i);
            const int dst = b * ldm + (i - k) * R + k;
            const int p0 = 1 + g * PP;
            if (g == 0) buf[lidx(dst)] = y0[c];
#pragma unroll
            for (int j = 0; j < PP; ++j) {
                int p = p0 + j;
                if (p <= hR) {
                    buf[lidx(dst + p * Ns)] = ya[c][j];
                    buf[lidx(dst + (R - p) * Ns)] = yb[c][j];
                }
            }
        }
    }
    __syncthreads();
    return true;
}

// The same stage for R <= 47 with the weights tabulated per term:
// wq[(q - 1) GP + p - 1] = w_{(p q) mod R} (GP = G PP columns, zero past hR), so
// the PP weights a work item needs for term q are consecutive and come in
// 16-byte pairs, with no modular index arithmetic in the term loop.  One work
// item per thread; returns false when that does not fit.
template <int EPT>
__device__ bool fft_stage_oddsym_t(float2* buf, int ldm, int nlines, int M, int Ns, int R,
                                   const float2* __restrict__ tw, int dir) {
    constexpr int PP = 8, MAXR = 47, MAXH = (MAXR - 1) / 2, MAXGP = ((MAXH + PP - 1) / PP) * PP;
    const int nbf = M / R, hR = (R - 1) / 2;
    const int G = (hR + PP - 1) / PP, GP = G * PP;
    const int total = nlines * nbf * G;
    const int nt = blockDim.x;
    if (R > MAXR || (R & 1) == 0 || total > nt) return false;
    __shared__ __align__(16) float2 wq[MAXH * MAXGP];
    for (int e = threadIdx.x; e < hR * GP; e += nt) {
        const int q = e / GP + 1, p = e - (q - 1) * GP + 1;
        wq[e] = p <= hR ? twiddle(tw, ((p * q) % R) * nbf, dir) : make_float2(0.f, 0.f);
    }
    __syncthreads();
    const int tstep = M / (Ns * R);
    const int t = threadIdx.x;
    float2 ya[PP], yb[PP], y0 = make_float2(0.f, 0.f);
    int g = 0, dst = 0;
    if (t < total) {
        const FDiv dG(G), dnbf(nbf), dNs(Ns);
        const int bi = dG.quo(t);
        g = t - bi * G;
        const int b = dnbf.quo(bi), i = bi - b * nbf;
        const int k = dNs.rem(i);
        const int src = b * ldm + i;
        dst = b * ldm + (i - k) * R + k;
        const float2 x0 = buf[lidx(src)];
        float2 A[PP], B[PP];
#pragma unroll
        for (int j = 0; j < PP; ++j) { A[j] = x0; B[j] = make_float2(0.f, 0.f); }
        float2 S0 = x0;
        const int e1 = k * tstep;
        const float2 w1 = twiddle(tw, e1, dir);
        const float2 w1c = make_float2(w1.x, -w1.y);
        float2 wa = make_float2(1.f, 0.f), wb = make_float2(1.f, 0.f);
        const float4* wrow = reinterpret_cast<const float4*>(wq + g * PP);
        for (int q = 1; q <= hR; ++q) {
            float2 xa = buf[lidx(src + q * nbf)], xb = buf[lidx(src + (R - q) * nbf)];
            if (Ns > 1) {
                if (((q - 1) & 7) == 0) {
                    wa = twiddle(tw, (int)(((long long)q * e1) % M), dir);
                    wb = twiddle(tw, (int)(((long long)(R - q) * e1) % M), dir);
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
        y0 = S0;
#pragma unroll
        for (int j = 0; j < PP; ++j) {
            const float2 iB = make_float2(-B[j].y, B[j].x);
            ya[j] = cadd(A[j], iB);
            yb[j] = csub(A[j], iB);
        }
    }
    __syncthreads();
    if (t < total) {
        const int p0 = 1 + g * PP;
        if (g == 0) buf[lidx(dst)] = y0;
#pragma unroll
        for (int j = 0; j < PP; ++j) {
            const int p = p0 + j;
            if (p <= hR) {
                buf[lidx(dst + p * Ns)] = ya[j];
                buf[lidx(dst + (R - p) * Ns)] = yb[j];
            }
        }
    }
    __syncthreads();
    return true;
}

// Full in-place FFT (dir = -1 forward, +1 inverse, unnormalised) of nlines
// lines of length plan.M.  Ends with a barrier.
template <int EPT>
__device__ void fft_lds(float2* buf, int ldm, int nlines, const FftPlan& plan, const float2* __restrict__ tw, int dir) {
    int Ns = 1;
    const int M = plan.M;
    for (int s = 0; s < plan.nst; ++s) {
        int R = plan.radix[s];
        switch (R) {
            case 2: fft_stage<2, EPT>(buf, ldm, nlines, M, Ns, tw, dir); break;
            case 3: fft_stage<3, EPT>(buf, ldm, nlines, M, Ns, tw, dir); break;
            case 4: fft_stage<4, EPT>(buf, ldm, nlines, M, Ns, tw, dir); break;
            case 5: fft_stage<5, EPT>(buf, ldm, nlines, M, Ns, tw, dir); break;
            case 8: fft_stage<8, EPT>(buf, ldm, nlines, M, Ns, tw, dir); break;
            default:
                if (!fft_stage_oddsym_t<EPT>(buf, ldm, nlines, M, Ns, R, tw, dir) &&
                    !fft_stage_oddsym<EPT>(buf, ldm, nlines, M, Ns, R, tw, dir))
                    fft_stage_generic<EPT>(buf, ldm, nlines, M, Ns, R, tw, dir);
                break;
        }
        Ns *= R;
    }
}

// Real-to-complex post-processing.  Line b holds Z = FFT_h(z) with
// z[n] = x[2n] + i x[2n+1] (N = 2h real samples).  Overwrites it with
// X[0..h-1] of the length-N real DFT, X[0] (real) and X[h] (real) packed as
// X[0] + i X[h] in slot 0.  twN[k] = exp(-2 pi i k / N).  idx(b, k) is the
// LDS slot of element k of line b.
template <class Idx>
__device__ inline void r2c_post_at(float2* buf, int nlines, int h, const float2* __restrict__ twN, Idx idx) {
    const int npairs = h / 2 + 1;
    const FDiv dnp(npairs);
    for (int t = threadIdx.x; t < nlines * npairs; t += blockDim.x) {
        int b = dnp.quo(t), k = t - b * npairs;
        if (k == 0) {
            float2 z = buf[idx(b, 0)];
            buf[idx(b, 0)] = make_float2(z.x + z.y, z.x - z.y);
        } else {
            int kk = h - k;
            float2 A = buf[idx(b, k)], B = buf[idx(b, kk)];
            float2 Bc = cconj(B);
            float2 E = make_float2(0.5f * (A.x + Bc.x), 0.5f * (A.y + Bc.y));
            float2 D = make_float2(0.5f * (A.x - Bc.x), 0.5f * (A.y - Bc.y));
            float2 O = make_float2(D.y, -D.x);  // D / i
            float2 WO = cmul(twN[k], O);
            buf[idx(b, k)] = cadd(E, WO);
            if (kk != k) buf[idx(b, kk)] = cconj(csub(E, WO));
        }
    }
    __syncthreads();
}

// Inverse of r2c_post: turns the packed half spectrum X into Z' such that the
// unnormalised inverse FFT_h of Z' gives N * x (even samples in .x, odd in .y).
template <class Idx>
__device__ inline void c2r_pre_at(float2* buf, int nlines, int h, const float2* __restrict__ twN, Idx idx) {
    const int npairs = h / 2 + 1;
    const FDiv dnp(npairs);
    for (int t = threadIdx.x; t < nlines * npairs; t += blockDim.x) {
        int b = dnp.quo(t), k = t - b * npairs;
        if (k == 0) {
            float2 p = buf[idx(b, 0)];  // X0 + i Xh
            buf[idx(b, 0)] = make_float2(p.x + p.y, p.x - p.y);
        } else {
            int kk = h - k;
            float2 P = buf[idx(b, k)], Q = buf[idx(b, kk)];
            float2 S = cadd(P, cconj(Q));
            float2 D = csub(P, cconj(Q));
            float2 w = twN[k];
            float2 iwc = make_float2(w.y, w.x);  // i * conj(w)
            buf[idx(b, k)] = cadd(S, cmul(iwc, D));
            if (kk != k) {
                // Z'[h-k] = conj(S) + i conj(W^{h-k}) (Q - conj P), conj(W^{h-k}) = -W^k
                float2 S2 = cconj(S);
                float2 D2 = cconj(make_float2(-D.x, -D.y));  // Q - conj(P) = -conj(D)
                float2 imw = make_float2(w.y, -w.x);          // -i * w
                buf[idx(b, kk)] = cadd(S2, cmul(imw, D2));
            }
        }
    }
    __syncthreads();
}

// lines at lidx(b * ldm + k)
__device__ inline void r2c_post(float2* buf, int ldm, int nlines, int h, const float2* __restrict__ twN) {
    r2c_post_at(buf, nlines, h, twN, [=](int b, int k) { return lidx(b * ldm + k); });
}
__device__ inline void c2r_pre(float2* buf, int ldm, int nlines, int h, const float2* __restrict__ twN) {
    c2r_pre_at(buf, nlines, h, twN, [=](int b, int k) { return lidx(b * ldm + k); });
}

// ---------------------------------------------------------------------------
// Register-resident radix-16 DFT (16 = 4 x 4 Cooley-Tukey), natural order in
// and out.  dir = -1 forward, +1 inverse.
__device__ __forceinline__ void dft16(float2* v, int dir) {
    float2 a[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        float2 t[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
        dft_small<4>(t, dir);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) a[n2][k1] = t[k1];
    }
    // internal twiddles W16^(n2 k1), W16 = exp(dir 2 pi i / 16)
    const float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f, r2 = 0.70710678118654752440f;
    const float d = (float)dir;
    a[1][1] = cmul(a[1][1], make_float2(c1, d * s1));
    a[1][2] = cmul(a[1][2], make_float2(r2, d * r2));
    a[1][3] = cmul(a[1][3], make_float2(s1, d * c1));
    a[2][1] = cmul(a[2][1], make_float2(r2, d * r2));
    a[2][2] = rot90(a[2][2], dir);
    a[2][3] = cmul(a[2][3], make_float2(-r2, d * r2));
    a[3][1] = cmul(a[3][1], make_float2(s1, d * c1));
    a[3][2] = cmul(a[3][2], make_float2(-r2, d * r2));
    a[3][3] = cmul(a[3][3], make_float2(-c1, -d * s1));
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        float2 t[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
        dft_small<4>(t, dir);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
    }
}

// twiddle powers w^q, q = 1..R-1, from table reads of w^(2^j) and products
template <int R>
__device__ __forceinline__ void twiddle_pows(const float2* __restrict__ tw, int base, int dir, float2* w) {
    w[1] = twiddle(tw, base, dir);
#pragma unroll
    for (int q = 2; q < R; ++q) {
        if ((q & (q - 1)) == 0) w[q] = twiddle(tw, q * base, dir);        // power of two: table
        else w[q] = cmul(w[q & (q - 1)], w[q & -q]);                        // product of two table values
    }
}
