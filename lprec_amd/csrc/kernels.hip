// Device kernels of the log-polar Radon pipeline for gfx950.
//
// Data layout in HBM (per handle, nz slices):
//   fc  [nz][N][N]          prefiltered, centred image (fwd input side)
//   Rc  [nz][P][N]          prefiltered, stuffed/filtered sinogram (adj input side)
//   S   [3][nz][h][Nrho]    theta half-spectra, stored transposed (rho contiguous)
//   G   [3][nz][Nrho][Ntheta] convolved log-polar grids, one per span
// The reference's cuFFT 2-D R2C/C2R (lpRgpu.cu:270-277, 314-321) becomes three
// passes: a row pass (gather + real FFT along theta), a column pass (FFT along
// rho, x kernel spectrum, inverse) and an inverse row pass; the texture
// gathers (main_kernels.cuh:30-84) become exact 16-tap B-spline sums from HBM
// with periodic addressing.
#include <algorithm>
#include <numeric>

#include "kernels.hpp"

namespace lpk {

// Logical index of hardware block b of a 1-D launch of n blocks.  The
// dispatcher deals workgroups to the 8 XCDs round-robin (b -> XCD b % 8); this
// gives each XCD a contiguous range of logical blocks, so logical neighbours
// (blocks that share source data) run concurrently behind the same L2.
__device__ __forceinline__ int xcd_block(int b, int n) {
    constexpr int X = 8;
    const int q = n / X, r = n - q * X;
    const int x = b % X, s = b / X;
    return x * q + min(x, r) + s;
}

__device__ __forceinline__ int wrapi(int i, int n) {
    if ((unsigned)i < (unsigned)n) return i;
    i %= n;
    return i < 0 ? i + n : i;
}

// kernel-spectrum entry, conjugated for the transposed convolution
__device__ __forceinline__ float2 fzval(const float2* __restrict__ fz, int i, int conj) {
    float2 v = fz[i];
    return conj ? make_float2(v.x, -v.y) : v;
}

// bspline_weights (main_kernels.cuh:9-19)
__device__ __forceinline__ void bspline_w(float f, float* w) {
    float one = 1.0f - f, sq = f * f, osq = one * one;
    w[0] = (1.0f / 6.0f) * osq * one;
    w[1] = (2.0f / 3.0f) - 0.5f * sq * (2.0f - f);
    w[2] = (2.0f / 3.0f) - 0.5f * osq * (2.0f - one);
    w[3] = (1.0f / 6.0f) * sq * f;
}

// Sample the periodic grid T[H][W] at table coordinate (x, y): the value the
// reference's cubicTex2D / linearTex2D return for u = x + 0.5, v = y + 0.5,
// with exact (not 8-bit fixed point) weights.
__device__ __forceinline__ float sample(const float* __restrict__ T, int W, int H, float x, float y, bool cubic) {
    float fx = floorf(x), fy = floorf(y);
    float ax = x - fx, ay = y - fy;
    int ix = (int)fx, iy = (int)fy;
    if (cubic) {
        float wx[4], wy[4];
        bspline_w(ax, wx);
        bspline_w(ay, wy);
        int c0, c1, c2, c3;
        if (ix >= 1 && ix < W - 2) {
            c0 = ix - 1; c1 = ix; c2 = ix + 1; c3 = ix + 2;
        } else {
            c0 = wrapi(ix - 1, W); c1 = wrapi(ix, W); c2 = wrapi(ix + 1, W); c3 = wrapi(ix + 2, W);
        }
        float acc = 0.f;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const float* row = T + (size_t)wrapi(iy - 1 + b, H) * W;
            float r = wx[0] * row[c0] + wx[1] * row[c1] + wx[2] * row[c2] + wx[3] * row[c3];
            acc += wy[b] * r;
        }
        return acc;
    }
    int c0 = wrapi(ix, W), c1 = wrapi(ix + 1, W);
    const float* r0 = T + (size_t)wrapi(iy, H) * W;
    const float* r1 = T + (size_t)wrapi(iy + 1, H) * W;
    return (1.f - ay) * ((1.f - ax) * r0[c0] + ax * r0[c1]) + ay * ((1.f - ax) * r1[c0] + ax * r1[c1]);
}

// Loads and stores at a block-uniform base plus a zero-extended 32-bit byte
// offset per lane (the base stays in scalar registers, no 64-bit lane math).
template <class T>
__device__ __forceinline__ T ld_at(const T* base, unsigned off) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off);
}
template <class T>
__device__ __forceinline__ void st_at(T* base, unsigned off, T v) {
    *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + off) = v;
}
// The same for float2 with the non-temporal hint (global_load/store ... nt):
// data that is streamed once and not re-read from the caches.  Measured on
// the inverse row pass (its S reads and G writes): 1.12 / 0.98 against
// 1.19 / 1.05 ms; on the column pass at Nrho = 4096 (S in through
// global_load_lds with the hint, kCpolNT, and S out): 1.26 / 1.28 against
// 1.33 / 1.35 ms, neither half alone faster; on the radix-16 inverse row
// pass at h = 512: 0.167 / 0.135 against 0.177 / 0.141 ms.  Slower on the
// forward row pass (F reads, S writes); no gain on k_cols16.
constexpr int kCpolNT = 2;   // cache-policy operand of the load-to-LDS builtin: nt
typedef float v2f_nt __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 ld_at_nt(const float2* base, unsigned off) {
    const v2f_nt v =
        __builtin_nontemporal_load(reinterpret_cast<const v2f_nt*>(reinterpret_cast<const char*>(base) + off));
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ void st_at_nt(float2* base, unsigned off, float2 w) {
    const v2f_nt v = {w.x, w.y};
    __builtin_nontemporal_store(v, reinterpret_cast<v2f_nt*>(reinterpret_cast<char*>(base) + off));
}

// Global -> LDS staging with 8 independent loads in flight per thread.
// load(i) returns element i (0 <= i < n), store(i, v) writes it to LDS.  The
// loads are unconditional (lanes past n reload element n - 1): a guarded load
// compiles to a branch and a wait of its own.
template <class Load, class Store>
__device__ __forceinline__ void stage_in(int n, Load load, Store store) {
    constexpr int U = 8;
    const int nt = blockDim.x;
    for (int i0 = threadIdx.x; i0 < n; i0 += U * nt) {
        float2 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = load(min(i0 + u * nt, n - 1));
#pragma unroll
        for (int u = 0; u < U; ++u) {
            int i = i0 + u * nt;
            if (i < n) store(i, v[u]);
        }
    }
}

// ---------------------------------------------------------------------------
// fwd input: copy3Dshifted (lpRgpu.cu:189) + x-direction prefilter (lpRgpu.cu:219-240)
__global__ void k_image_rows(const float* __restrict__ f, float* __restrict__ fc, int N0, int N, int o, int cubic) {
    extern __shared__ float sm[];
    float* line = sm;
    float* tmp = sm + N;
    const int yi = blockIdx.x, z = blockIdx.y;
    const bool rowin = yi >= o && yi < o + N0;
    const float* src = f + ((size_t)z * N0 + (rowin ? yi - o : 0)) * N0;
    // 8 unconditional loads in flight per lane (clamped into the row, zeroed
    // outside it), then the LDS stores
    for (int x0 = threadIdx.x; x0 < N; x0 += 8 * blockDim.x) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = src[min(max(x0 + u * (int)blockDim.x - o, 0), N0 - 1)];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int x = x0 + u * blockDim.x, xs = x - o;
            if (x < N) line[x] = (rowin && xs >= 0 && xs < N0) ? v[u] : 0.f;
        }
    }
    __syncthreads();
    if (cubic) iir_line_lds(line, tmp, N, threadIdx.x, blockDim.x);
    float* dst = fc + ((size_t)z * N + yi) * N;
    for (int x = threadIdx.x; x < N; x += blockDim.x) dst[x] = line[x];
}

// The FBP filter of one line of N reals (in place): the length-2Lh circular
// convolution (2Lh >= 2N) with the kernel restricted to lags
// |l| < N, whose spectrum H2 [Lh+1] is real; the N outputs equal the
// length-2N circular (and the reference's 4N padded) convolution.  The 2Lh
// real transform is a complex FFT of length Lh with R2C/C2R post-processing.
template <int EPT>
__device__ void filter_line(float2* smc, float* line, const SinoPrep& sp) {
    const int N = sp.N, Lh = sp.Lh;
    for (int e = threadIdx.x; e < Lh; e += blockDim.x)
        smc[lidx(e)] = 2 * e < N ? make_float2(line[2 * e], line[2 * e + 1]) : make_float2(0.f, 0.f);
    __syncthreads();
    fft_lds<EPT>(smc, Lh, 1, sp.planN, sp.twN, -1);
    r2c_post(smc, Lh, 1, Lh, sp.tw2N);
    for (int k = threadIdx.x; k < Lh; k += blockDim.x) {
        float2 v = smc[lidx(k)];
        if (k == 0) smc[lidx(0)] = make_float2(v.x * sp.H2[0], v.y * sp.H2[Lh]);
        else smc[lidx(k)] = make_float2(v.x * sp.H2[k], v.y * sp.H2[k]);
    }
    __syncthreads();
    c2r_pre(smc, Lh, 1, Lh, sp.tw2N);
    fft_lds<EPT>(smc, Lh, 1, sp.planN, sp.twN, +1);
    for (int e = threadIdx.x; e < N / 2; e += blockDim.x) {
        float2 v = smc[lidx(e)];
        line[2 * e] = v.x;
        line[2 * e + 1] = v.y;
    }
}

// adj input: copy3Dshifted (lpRgpu.cu:201) + stepangles (:202) + padker (:203) +
// applyFilter (:333-382) + s-direction prefilter.  The reference's FBP filter is a
// circular convolution of the 4N zero-padded row with a real even kernel; only lags
// |l| < N reach the N outputs, so it is evaluated exactly as a length-2N circular
// convolution with the kernel restricted to those lags (spectrum H2, built on host).
template <int EPT, bool FILT>
__global__ void k_sino_rows(const float* __restrict__ R, float* __restrict__ Rc, SinoPrep sp, int cubic) {
    extern __shared__ float2 smc[];
    const int N = sp.N;
    // with the cubic prefilter the grid covers the kept angles only (the
    // column sweep does not read the stuffed rows); without it, every row
    const int ai = cubic ? blockIdx.x * sp.os : blockIdx.x, z = blockIdx.y;
    float* dst = Rc + ((size_t)z * sp.P + ai) * N;
    if (ai % sp.os != 0) {   // stuffed row: zeros
        for (int x = threadIdx.x; x < N; x += blockDim.x) dst[x] = 0.f;
        return;
    }
    const int a = ai / sp.os;
    const float* src = R + ((size_t)z * sp.P0 + a) * sp.N0;
    // N + N floats after the FFT buffer (which is only allocated with a filter)
    float* line = FILT ? reinterpret_cast<float*>(smc + lds_slots(sp.Lh)) : reinterpret_cast<float*>(smc);
    float* tmp = line + N;
    // 8 loads in flight per lane, then the LDS stores
    for (int x0 = threadIdx.x; x0 < N; x0 += 8 * blockDim.x) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = src[min(max(x0 + u * (int)blockDim.x - sp.lb, 0), sp.N0 - 1)];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (x0 + u * (int)blockDim.x < N) line[x0 + u * blockDim.x] = v[u];
    }
    if (FILT) {
        __syncthreads();
        filter_line<EPT>(smc, line, sp);
    }
    __syncthreads();
    if (cubic) iir_line_lds(line, tmp, N, threadIdx.x, blockDim.x);
    for (int x = threadIdx.x; x < N; x += blockDim.x) dst[x] = line[x];
}

// The same preparation for two kept angles per block when a filter is set: the
// kernel's spectrum is real and even, so the complex line x0 + i x1 filtered
// by one length-2N complex FFT pair gives the two filtered lines as its real
// and imaginary parts (no R2C/C2R steps, half the FFT launches per line).
// Without the cubic prefilter the block also writes the zero (stuffed) rows
// that follow each angle; the cubic column sweep does not read them.
template <int EPT>
__global__ void k_sino_rows_pair(const float* __restrict__ R, float* __restrict__ Rc, SinoPrep sp, int cubic) {
    extern __shared__ float2 smc[];
    const int N = sp.N, M = 2 * N;
    const int a0 = 2 * blockIdx.x, z = blockIdx.y;
    const int nrow = min(2, sp.P0 - a0);
    float* line = reinterpret_cast<float*>(smc + lds_slots(M));   // [2][N]
    float* tmp = line + 2 * N;                                     // [2][N]
    for (int u = 0; u < (cubic ? 0 : nrow); ++u)
        for (int j = 1; j < sp.os; ++j) {
            float* dz = Rc + ((size_t)z * sp.P + (size_t)(a0 + u) * sp.os + j) * N;
            for (int x = threadIdx.x; x < N; x += blockDim.x) dz[x] = 0.f;
        }
    // both rows' loads in flight together, 4 per row per lane per round
    {
        const float* src0 = R + ((size_t)z * sp.P0 + a0) * sp.N0;
        const float* src1 = R + ((size_t)z * sp.P0 + min(a0 + 1, sp.P0 - 1)) * sp.N0;
        for (int x0 = threadIdx.x; x0 < N; x0 += 4 * blockDim.x) {
            float v0[4], v1[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int xs = min(max(x0 + u * (int)blockDim.x - sp.lb, 0), sp.N0 - 1);
                v0[u] = src0[xs];
                v1[u] = src1[xs];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int x = x0 + u * blockDim.x;
                if (x < N) {
                    line[x] = v0[u];
                    line[N + x] = nrow > 1 ? v1[u] : 0.f;
                }
            }
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < M; e += blockDim.x)
        smc[lidx(e)] = e < N ? make_float2(line[e], line[N + e]) : make_float2(0.f, 0.f);
    __syncthreads();
    fft_lds<EPT>(smc, M, 1, sp.plan2N, sp.tw2Nc, -1);
    for (int k = threadIdx.x; k < M; k += blockDim.x) {
        const float hk = sp.H2[k <= N ? k : M - k];
        float2 v = smc[lidx(k)];
        smc[lidx(k)] = make_float2(v.x * hk, v.y * hk);
    }
    __syncthreads();
    fft_lds<EPT>(smc, M, 1, sp.plan2N, sp.tw2Nc, +1);
    for (int e = threadIdx.x; e < N; e += blockDim.x) {
        float2 v = smc[lidx(e)];
        line[e] = v.x;
        line[N + e] = v.y;
    }
    __syncthreads();
    // the two lines' prefilters run side by side on the two halves of the block
    const int half = blockDim.x / 2, u = threadIdx.x < half ? 0 : 1;
    if (cubic) iir_line_lds(line + u * N, tmp + u * N, N, threadIdx.x - u * half, half);
    for (int v = 0; v < nrow; ++v) {
        float* dst = Rc + ((size_t)z * sp.P + (size_t)(a0 + v) * sp.os) * N;
        for (int x = threadIdx.x; x < N; x += blockDim.x) dst[x] = line[v * N + x];
    }
}

// y-direction prefilter (SamplesToCoefficients2DY after the transposes of
// prefilter2D, lpRgpu.cu:229-239): one thread per column, T rows per chunk with a
// kHalo-row halo on each side held in registers.
// Out of place: chunks read halos that neighbouring chunks write.
// stuff > 1: only rows that are multiples of stuff hold data (the angle
// stuffing of the sinogram); the others are zero and are not read.
template <int T>
__global__ void k_col_iir(const float* __restrict__ in, float* __restrict__ outp, int L, int W, int stuff) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int r0 = blockIdx.y * T;
    const int z = blockIdx.z;
    if (x >= W) return;
    constexpr int K = kHalo;
    constexpr int NV = T + 2 * K;
    const float* col = in + (size_t)z * L * W + x;
    float* ocol = outp + (size_t)z * L * W + x;
    // Interior chunks load without per-row conditions, so that all loads are
    // in flight together (a guarded load compiles to a branch and a wait of
    // its own); r0 - K is even, so with stuff 2 the data rows are the even n.
    float v[NV];
    const bool interior = r0 >= K && r0 + T + K <= L;
    const float* c0 = col + (size_t)(r0 - K) * W;
    if (interior && stuff == 1) {
#pragma unroll
        for (int n = 0; n < NV; ++n) v[n] = c0[(size_t)n * W];
    } else if (interior && stuff == 2) {
#pragma unroll
        for (int n = 0; n < NV; ++n) v[n] = (n & 1) ? 0.f : c0[(size_t)n * W];
    } else {
#pragma unroll
        for (int n = 0; n < NV; ++n) {
            int row = r0 - K + n;   // uniform over the block
            v[n] = (row >= 0 && row < L && row % stuff == 0) ? col[(size_t)row * W] : 0.f;
        }
    }
    const float zp = kPole;
    if (r0 == 0) {
        float sum = v[K];
        float zn = zp;
#pragma unroll
        for (int n = 0; n < kHorizon; ++n) {
            if (n < L) sum += zn * v[K + n];
            zn *= zp;
        }
        float c = kLambda * sum;
        v[K] = c;
#pragma unroll
        for (int n = K + 1; n < NV; ++n) { c = kLambda * v[n] + zp * c; v[n] = c; }
    } else {
        float c = kLambda * v[0];
        v[0] = c;
#pragma unroll
        for (int n = 1; n < NV; ++n) { c = kLambda * v[n] + zp * c; v[n] = c; }
    }
    const int last = min(r0 + T + K - 1, L - 1) - (r0 - K);
    const bool exact_end = (r0 + T + K - 1) >= (L - 1);
    float c = 0.f;
#pragma unroll
    for (int n = NV - 1; n >= 0; --n) {
        if (n > last) continue;
        if (n == last) c = exact_end ? (zp / (zp - 1.0f)) * v[n] : -zp * v[n];
        else c = zp * (c - v[n]);
        v[n] = c;
    }
#pragma unroll
    for (int n = K; n < K + T; ++n) {
        int row = r0 - K + n;
        if (row < L) ocol[(size_t)row * W] = v[n];
    }
}

// ---------------------------------------------------------------------------
// Row pass, forward direction: for R log-polar rows of one span, load the
// gathered samples F (interp + mulexp of lpRgpu.cu:263-267 / 307-311, done by
// k_gather), real FFT along theta, store the half spectra transposed into S.
template <int EPT>
__global__ void k_rows_fwd(RowPass rp, const float* __restrict__ F, const int2* __restrict__ range, int ldm) {
    extern __shared__ float2 smc[];
    const int Nt = rp.Ntheta, h = rp.h, R = rp.R, Nr = rp.Nrho;
    const int row0 = blockIdx.x * R, span = blockIdx.y;
    const int z0 = blockIdx.z * rp.zgroup;
    const int z1 = min(z0 + rp.zgroup, rp.nz);
    // sampled theta range of each of the block's rows (the host keeps R <= 8
    // for these kernels)
    __shared__ int2 rng_s[8];
    if (threadIdx.x < R) rng_s[threadIdx.x] = range ? range[(size_t)span * Nr + row0 + threadIdx.x] : make_int2(0, Nt);
    __syncthreads();
    const FDiv dh(h), dR(R);
    for (int z = z0; z < z1; ++z) {
        const float* Fz = F + (((size_t)span * rp.nz + z) * Nr + row0) * (size_t)Nt;
        stage_in(
            R * h,
            [&](int p) {
                // unconditional load at an index clamped into the row's range
                // (its lines are loaded anyway); pairs outside it are zeroed
                const int r = dh.quo(p), th = 2 * (p - r * h);
                const int2 lr = rng_s[r];
                const int tlo = min(max(lr.x & ~1, 0), Nt - 2), thi = max(tlo, min((lr.y - 1) & ~1, Nt - 2));
                float2 v = ld_at(reinterpret_cast<const float2*>(Fz),
                                 (unsigned)(r * Nt + min(max(th, tlo), thi)) * 4u);
                v.x = (th >= lr.x && th < lr.y) ? v.x : 0.f;
                v.y = (th + 1 >= lr.x && th + 1 < lr.y) ? v.y : 0.f;
                return v;
            },
            [&](int p, float2 v) {
                int r = dh.quo(p), e = p - r * h;
                smc[lidx(r * ldm + e)] = v;
            });
        __syncthreads();
        fft_lds<EPT>(smc, ldm, R, rp.planh, rp.twh, -1);
        r2c_post(smc, ldm, R, h, rp.twN);
        float2* out = rp.S + ((size_t)span * rp.nz + z) * h * (size_t)Nr + row0;
        for (int q = threadIdx.x; q < R * h; q += blockDim.x) {
            int f = dR.quo(q), r = q - f * R;
            out[(size_t)f * Nr + r] = smc[lidx(r * ldm + f)];
        }
        __syncthreads();
    }
}

// LDS-tiled gather (the interp kernel of main_kernels.cuh:64-84 for all four
// grids): the block stages the source box of its output tile with coalesced
// row loads, then evaluates every sample of the tile from LDS.  A block owns
// GROUP tiles that share one output tile (the three spans of the
// LP -> Cartesian gather, lpRgpu.cu:328, summed in span order); the
// single-span gathers run k_gather1, which keeps each point's sample
// geometry across the slices of its group.
// KIND only names the instance (0 image->LP, 1 sinogram->LP, 2 LP->polar,
// 3 LP->Cartesian) so that profiles separate the four gathers.


// Lane (lr, lc) of a gather tile with nc = 2^lnc columns: the lane owns the
// points of rows lr + e * rstep (rstep = kGatherThreads / nc), column lc.
// The lanes of a wave form a block of up to 16 columns by 64 / 16 rows (one
// row of 64 columns would spread its 16-tap reads along an arc of the source
// box and over few rows of LDS banks): bits 0-3 of the thread index give the
// column within a 16-wide band, the next log2(rstep) bits the row, the rest
// the band.
__device__ __forceinline__ void gather_lane(int lnc, int& lr, int& lc) {
    const int t = threadIdx.x;
    const int lbw = min(lnc, 4), lrs = __builtin_ctz((unsigned)kGatherThreads) - lnc;
    lc = (t & ((1 << lbw) - 1)) | ((t >> (lbw + lrs)) << lbw);
    lr = (t >> lbw) & ((1 << lrs) - 1);
}

// Row and column of point p of a tile with nc columns, from a float
// reciprocal (exact for p, nc < 2^20) instead of an integer division.
__device__ __forceinline__ void tile_rc(int p, int nc, float inv_nc, int& r, int& c) {
    r = (int)(((float)p + 0.5f) * inv_nc);
    c = p - r * nc;
}

// Stage the box [x0, x0+bw) x [y0, y0+bh) of the periodic grid S [H][W] into
// LDS rows of stride bs (>= bw, chosen on the host), UL loads in flight per
// lane.  A box inside the grid (nearly all of them) needs no wrapping: element
// i = yy bw + xx sits at source offset i + yy (W - bw) and LDS offset
// i + yy (bs - bw).
template <int UL>
__device__ __forceinline__ void load_box(float* box, const float* __restrict__ S, const GTile& t, int W, int H) {
    constexpr int nt = kGatherThreads;
    const int nbox = t.bw * t.bh;
    const int bs = t.bs;
    // (i + 0.5) / bw in float is exact enough for i, bw < 2^20
    const float inv = 1.0f / (float)t.bw;
    if (t.x0 >= 0 && t.y0 >= 0 && t.x0 + t.bw <= W && t.y0 + t.bh <= H) {
        const float* Sb = S + (size_t)t.y0 * W + t.x0;
        const int sw = W - t.bw, lw = bs - t.bw;
        for (int i0 = threadIdx.x; i0 < nbox; i0 += UL * nt) {
            float v[UL];
            int li[UL];
#pragma unroll
            for (int u = 0; u < UL; ++u) {
                const int i = i0 + u * nt;
                v[u] = 0.f;
                li[u] = 0;
                if (i < nbox) {
                    const int yy = (int)(((float)i + 0.5f) * inv);
                    v[u] = ld_at(Sb, (unsigned)(i + yy * sw) * 4u);
                    li[u] = i + yy * lw;
                }
            }
#pragma unroll
            for (int u = 0; u < UL; ++u)
                if (i0 + u * nt < nbox) box[li[u]] = v[u];
        }
        return;
    }
    for (int i0 = threadIdx.x; i0 < nbox; i0 += UL * nt) {
        float v[UL];
#pragma unroll
        for (int u = 0; u < UL; ++u) {
            int i = i0 + u * nt;
            v[u] = 0.f;
            if (i < nbox) {
                int yy = (int)(((float)i + 0.5f) * inv);
                int xx = i - yy * t.bw;
                unsigned off = (unsigned)wrapi(t.y0 + yy, H) * (unsigned)W + (unsigned)wrapi(t.x0 + xx, W);
                v[u] = S[off];
            }
        }
#pragma unroll
        for (int u = 0; u < UL; ++u) {
            int i = i0 + u * nt;
            if (i < nbox) {
                int yy = (int)(((float)i + 0.5f) * inv);
                box[i + yy * (bs - t.bw)] = v[u];
            }
        }
    }
}

// Cubic B-spline weights times 6: (1-f)^3, 3f^3 - 6f^2 + 4, -3f^3 + 3f^2 +
// 3f + 1, f^3 (the 1/36 of the two axes is applied once per sample).
__device__ __forceinline__ void bspline_w6(float f, float* w) {
    const float one = 1.0f - f, sq = f * f;
    w[0] = one * one * one;
    w[1] = fmaf(sq, fmaf(3.0f, f, -6.0f), 4.0f);
    w[2] = fmaf(f, fmaf(f, fmaf(-3.0f, f, 3.0f), 3.0f), 1.0f);
    w[3] = sq * f;
}

// One sample of the staged box (row stride bs) at source coordinate c (valid).
// The LDS slot of tap (0, 0) is formed in float from box-relative integers
// (exact below 2^24) and clamped into the box in one step; the tile boxes hold
// every tap of their points, so the clamp only guards the LDS bounds.
__device__ __forceinline__ float box_sample(const float* box, const GTile& t, float2 c, int cubic) {
    const float fx = floorf(c.x), fy = floorf(c.y);
    const float ax = c.x - fx, ay = c.y - fy;
    const int sy = t.bs;
    if (cubic) {
        float slot = fmaf(fy - (float)(t.y0 + 1), (float)sy, fx - (float)(t.x0 + 1));
        slot = __builtin_amdgcn_fmed3f(slot, 0.f, (float)((t.bh - 4) * sy + t.bw - 4));
        const float* q = box + (int)slot;
        float wx[4], wy[4];
        bspline_w6(ax, wx);
        bspline_w6(ay, wy);
        float v = 0.f;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const float* qr = q + b * sy;
            v += wy[b] * (wx[0] * qr[0] + wx[1] * qr[1] + wx[2] * qr[2] + wx[3] * qr[3]);
        }
        return v * (1.0f / 36.0f);
    }
    float slot = fmaf(fy - (float)t.y0, (float)sy, fx - (float)t.x0);
    slot = __builtin_amdgcn_fmed3f(slot, 0.f, (float)((t.bh - 2) * sy + t.bw - 2));
    const float* q1 = box + (int)slot;
    const float* q2 = q1 + sy;
    return (1.f - ay) * ((1.f - ax) * q1[0] + ax * q1[1]) + ay * ((1.f - ax) * q2[0] + ax * q2[1]);
}

// Stage the box of tile t like load_box, but, when the padded box (rows of
// stride bs including the pad columns) lies inside the grid, with
// loads that write LDS directly, 4 bytes per lane: 64 consecutive box floats
// are 64 consecutive LDS floats, no registers or LDS stores in between.  The
// last chunk's lanes past the box reload its origin into LDS floats below the
// next multiple of 64 (the launch rounds the box allocation up to that).
__device__ __forceinline__ void stage_box(float* box, const float* __restrict__ S, const GTile& t, int W, int H) {
    const int bs = t.bs, nb = bs * t.bh;
    if (t.x0 >= 0 && t.y0 >= 0 && t.x0 + bs <= W && t.y0 + t.bh <= H) {
        const float* Sb = S + (size_t)t.y0 * W + t.x0;
        const float inv_bs = 1.0f / (float)bs;
        const int lane = threadIdx.x & 63;
        for (int k = (threadIdx.x >> 6) * 64; k < nb; k += kGatherThreads) {
            const int i = k + lane;
            const int yy = (int)(((float)i + 0.5f) * inv_bs);
            const float* src = Sb + (i < nb ? yy * W + (i - yy * bs) : 0);
            __builtin_amdgcn_global_load_lds(src, box + k, 4, 0, 0);
        }
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    } else {
        load_box<8>(box, S, t, W, H);
    }
}

// Single-span gathers with the sample geometry cached: the LDS slot of tap
// (0,0), the fractional offsets and the output scale of the block's points are
// computed once and reused for every slice of the block's group (only the
// staged box changes from slice to slice).
template <int KIND>
__global__ void __launch_bounds__(kGatherThreads, 5) k_gather1(Gather g, const GTile* __restrict__ tiles) {
    extern __shared__ float box[];
    constexpr int PPT = kGatherPPT;
    const int nzg = (g.nz + g.zgroup - 1) / g.zgroup;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int zg = lb % nzg, tile = lb / nzg;
    const GTile t = tiles[tile];
    int lr, lc;
    gather_lane(t.lnc, lr, lc);
    const int rstep = kGatherThreads >> t.lnc;
    const unsigned o_lane = (unsigned)(lr * g.o_row + lc) * 4u;
    const unsigned c_lane = (unsigned)(lr * g.c_row + lc) * 8u;
    const unsigned s_lane = (unsigned)(lr * g.s_row + lc) * 4u;
    const int z0 = zg * g.zgroup, z1 = min(z0 + g.zgroup, g.nz);
    const float2* cb = g.coord + (size_t)t.span * g.c_span + (size_t)(t.r0 + g.c_ro) * g.c_row + t.c0 + g.c_co;
    const float* sb = g.scale ? g.scale + (size_t)t.r0 * g.s_row + t.c0 : nullptr;
    const int sy = t.bs;
    const int cubic = g.cubic;
    // box-relative origin of a sample's first tap and the largest valid slot
    const float ox = (float)(t.x0 + (cubic ? 1 : 0)), oy = (float)(t.y0 + (cubic ? 1 : 0));
    const float maxslot = cubic ? (float)((t.bh - 4) * sy + t.bw - 4) : (float)((t.bh - 2) * sy + t.bw - 2);
    int slot[PPT];
    float ax[PPT], ay[PPT], sc[PPT];
    float2 c[PPT];
#pragma unroll
    for (int e = 0; e < PPT; ++e) {
        c[e] = make_float2(__int_as_float(0x7fc00000), 0.f);
        sc[e] = 1.f;
        if (lr + e * rstep < t.nr) {
            c[e] = ld_at(cb + (size_t)(e * rstep) * g.c_row, c_lane);
            if (sb) sc[e] = ld_at(sb + (size_t)(e * rstep) * g.s_row, s_lane);
        }
    }
#pragma unroll
    for (int e = 0; e < PPT; ++e) {
        const float fx = floorf(c[e].x), fy = floorf(c[e].y);
        ax[e] = c[e].x - fx;
        ay[e] = c[e].y - fy;
        const float sl = fmaf(fy - oy, (float)sy, fx - ox);
        slot[e] = c[e].x == c[e].x ? (int)__builtin_amdgcn_fmed3f(sl, 0.f, maxslot) : -1;
    }
    for (int z = z0; z < z1; ++z) {
        float* O = g.out + (size_t)z * g.o_z + (size_t)t.span * g.o_span + (size_t)t.r0 * g.o_row + t.c0;
        if (t.bw > 0) {
            const float* S = g.src + (size_t)z * g.src_z + (size_t)t.span * g.src_span;
            // the gathers into the LP grids (from the image or the sinogram)
            // stage directly; the polar output's boxes measured faster
            // through registers
            if (KIND == 2) load_box<8>(box, S, t, g.W, g.H);
            else stage_box(box, S, t, g.W, g.H);
        }
        __syncthreads();
        // two points per step; the cached values of the step are selected by
        // uniform branches so that the arrays stay in registers
#pragma unroll 1
        for (int e = 0; e < PPT; e += 2) {
            int sl[2];
            float fx[2], fy[2], s2[2];
#pragma unroll
            for (int q = 0; q < PPT; q += 2)
                if (q == e) {
#pragma unroll
                    for (int u = 0; u < 2; ++u) { sl[u] = slot[q + u]; fx[u] = ax[q + u]; fy[u] = ay[q + u]; s2[u] = sc[q + u]; }
                }
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                if (lr + (e + u) * rstep < t.nr) {
                    float v = 0.f;
                    if (sl[u] >= 0) {
                        const float* q = box + sl[u];
                        if (cubic) {
                            float wx[4], wy[4];
                            bspline_w6(fx[u], wx);
                            bspline_w6(fy[u], wy);
#pragma unroll
                            for (int b = 0; b < 4; ++b) {
                                const float* qr = q + b * sy;
                                v += wy[b] * (wx[0] * qr[0] + wx[1] * qr[1] + wx[2] * qr[2] + wx[3] * qr[3]);
                            }
                            v *= (1.0f / 36.0f);
                        } else {
                            const float* q2 = q + sy;
                            v = (1.f - fy[u]) * ((1.f - fx[u]) * q[0] + fx[u] * q[1]) + fy[u] * ((1.f - fx[u]) * q2[0] + fx[u] * q2[1]);
                        }
                    }
                    st_at(O + (size_t)((e + u) * rstep) * g.o_row, o_lane, v * s2[u]);
                }
            }
        }
        __syncthreads();
    }
}

// Image -> log-polar samples taken straight from the prefiltered image in
// global memory (16 taps per point through the caches), one thread per point
// and no LDS box: the sampling that a row pass fused with this gather would
// have to do (its rows' source arcs do not fit in LDS).  Kept to measure that
// design against the LDS-tiled k_gather1<0>; only the sampled theta range of
// each row is written, as by the tiles.  Grid (Ntheta / 256, Nrho, 3 nz).
__global__ void __launch_bounds__(256) k_gather_direct(Gather g, const int2* __restrict__ range, int Nr, int Nt) {
    const int th = blockIdx.x * 256 + threadIdx.x;
    const int row = blockIdx.y;
    const int span = blockIdx.z / g.nz, z = blockIdx.z - span * g.nz;
    if (th >= Nt) return;
    const int2 rg = range ? range[(size_t)span * Nr + row] : make_int2(0, Nt);
    if (th < rg.x || th >= rg.y) return;
    const float2 c = g.coord[(size_t)span * g.c_span + (size_t)row * g.c_row + th];
    float v = 0.f;
    if (c.x == c.x) v = sample(g.src + (size_t)z * g.src_z + (size_t)span * g.src_span, g.W, g.H, c.x, c.y, g.cubic != 0);
    if (g.scale) v *= g.scale[(size_t)row * g.s_row + th];
    g.out[(size_t)z * g.o_z + (size_t)span * g.o_span + (size_t)row * g.o_row + th] = v;
}

void launch_gather_direct(const Gather& g, const int2* range, int Nr, int Nt, hipStream_t s) {
    hipLaunchKernelGGL(k_gather_direct, dim3((Nt + 255) / 256, Nr, 3 * g.nz), dim3(256), 0, s, g, range, Nr, Nt);
}

template <int KIND, int GROUP>
__global__ void __launch_bounds__(kGatherThreads) k_gather(Gather g, const GTile* __restrict__ tiles) {
    static_assert(GROUP > 1, "single-span gathers run k_gather1");
    extern __shared__ float box[];
    constexpr int PPT = kGatherPPT;
    // coordinates run GP points ahead of the sampling; the first GP are
    // issued before the box load and arrive while it streams in
    constexpr int GP = kGatherPipe;
    static_assert(PPT % GP == 0, "points per thread come in groups of GP");
    // logical order: the slice groups of one tile are neighbours (they share
    // the coordinates), then neighbouring tiles (they share source rows)
    const int nzg = (g.nz + g.zgroup - 1) / g.zgroup;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int zg = lb % nzg, tile = lb / nzg;
    const GTile* tb = tiles + (size_t)tile * GROUP;
    const GTile t0 = tb[0];
    // the tile has nc = 2^lnc columns: lane (lr, lc) holds the points of rows
    // lr + e * rstep, column lc, so every per-point address is a uniform base
    // (row e * rstep) plus a fixed lane offset
    int lr, lc;
    gather_lane(t0.lnc, lr, lc);
    const int rstep = kGatherThreads >> t0.lnc;
    const unsigned o_lane = (unsigned)(lr * g.o_row + lc) * 4u;
    const unsigned c_lane = (unsigned)(lr * g.c_row + lc) * 8u;
    const unsigned s_lane = (unsigned)(lr * g.s_row + lc) * 4u;
    const int z0 = zg * g.zgroup, z1 = min(z0 + g.zgroup, g.nz);
    const float* sb = g.scale ? g.scale + (size_t)t0.r0 * g.s_row + t0.c0 : nullptr;
    for (int z = z0; z < z1; ++z) {
        float* O = g.out + (size_t)z * g.o_z + (size_t)t0.span * g.o_span + (size_t)t0.r0 * g.o_row + t0.c0;
        float acc[PPT];
#pragma unroll
        for (int e = 0; e < PPT; ++e) acc[e] = 0.f;
#pragma unroll 1
        for (int gi = 0; gi < GROUP; ++gi) {
            const GTile t = tb[gi];
            const float2* cb = g.coord + (size_t)t.span * g.c_span + (size_t)(t.r0 + g.c_ro) * g.c_row + t.c0 + g.c_co;
            auto coord_of = [&](int e) {
                float2 c = make_float2(__int_as_float(0x7fc00000), 0.f);   // NaN: no sample
                if (lr + e * rstep < t0.nr) c = ld_at(cb + (size_t)(e * rstep) * g.c_row, c_lane);
                return c;
            };
            float2 cur[GP];
#pragma unroll
            for (int u = 0; u < GP; ++u) cur[u] = coord_of(u);
            if (t.bw > 0) load_box<8>(box, g.src + (size_t)z * g.src_z + (size_t)t.span * g.src_span, t, g.W, g.H);
            __syncthreads();
#pragma unroll 1
            for (int e = 0; e < PPT; e += GP) {
                float2 nxt[GP];
#pragma unroll
                for (int u = 0; u < GP; ++u) nxt[u] = cur[u];
                if (e + GP < PPT) {
#pragma unroll
                    for (int u = 0; u < GP; ++u) nxt[u] = coord_of(e + GP + u);
                }
                float vals[GP];
#pragma unroll
                for (int u = 0; u < GP; ++u) {
                    const bool in = lr + (e + u) * rstep < t0.nr;
                    const float2 c = cur[u];
                    vals[u] = (in && c.x == c.x) ? box_sample(box, t, c, g.cubic) : 0.f;
                }
                // acc is indexed by a loop variable: keep it in registers by
                // selecting the group explicitly
#pragma unroll
                for (int q = 0; q < PPT; q += GP)
                    if (q == e) {
#pragma unroll
                        for (int u = 0; u < GP; ++u) acc[q + u] += vals[u];
                    }
#pragma unroll
                for (int u = 0; u < GP; ++u) cur[u] = nxt[u];
            }
            __syncthreads();
        }
#pragma unroll
        for (int e = 0; e < PPT; ++e) {
            if (lr + e * rstep < t0.nr) {
                float v = acc[e];
                if (sb) v *= ld_at(sb + (size_t)(e * rstep) * g.s_row, s_lane);
                st_at(O + (size_t)(e * rstep) * g.o_row, o_lane, v);
            }
        }
    }
}

// Row pass, inverse direction: C2R along theta for R rows, writing G.
template <int EPT>
__global__ void k_rows_inv(RowPass rp, float* __restrict__ G, const int2* __restrict__ wrange, int ldm) {
    extern __shared__ float2 smc[];
    const int Nt = rp.Ntheta, h = rp.h, R = rp.R, Nr = rp.Nrho;
    const int row0 = blockIdx.x * R, span = blockIdx.y;
    const int z0 = blockIdx.z * rp.zgroup;
    const int z1 = min(z0 + rp.zgroup, rp.nz);
    const FDiv dh(h), dR(R);
    for (int z = z0; z < z1; ++z) {
        const float2* in = rp.S + ((size_t)span * rp.nz + z) * h * (size_t)Nr + row0;
        stage_in(
            R * h,
            [&](int q) {
                int f = dR.quo(q), r = q - f * R;
                return in[(size_t)f * Nr + r];
            },
            [&](int q, float2 v) {
                int f = dR.quo(q), r = q - f * R;
                smc[lidx(r * ldm + f)] = v;
            });
        __syncthreads();
        c2r_pre(smc, ldm, R, h, rp.twN);
        fft_lds<EPT>(smc, ldm, R, rp.planh, rp.twh, +1);
        float* out = G + (((size_t)span * rp.nz + z) * Nr + row0) * (size_t)Nt;
        for (int p = threadIdx.x; p < R * h; p += blockDim.x) {
            int r = dh.quo(p), e = p - r * h;
            int2 wr = wrange ? wrange[(size_t)span * Nr + row0 + r] : make_int2(0, Nt);
            if (2 * e + 1 >= wr.x && 2 * e < wr.y)
                *reinterpret_cast<float2*>(out + (size_t)r * Nt + 2 * e) = smc[lidx(r * ldm + e)];
        }
        __syncthreads();
    }
}

// Column pass: FFT along rho of one theta-frequency column, multiply by the
// kernel spectrum (mul, simple_kernels.cuh:2-12, scale folded into fzT), inverse
// FFT.  Column 0 carries the theta-DC and theta-Nyquist columns packed as
// X0 + i Xh; they are separated by Hermitian symmetry, multiplied by their own
// spectra and re-packed.
template <int EPT>
__global__ void k_cols(RowPass rp, FftPlan planM, const float2* __restrict__ twM, const float2* __restrict__ fzT) {
    extern __shared__ float2 smc[];
    const int M = rp.Nrho, h = rp.h;
    const int f = blockIdx.x, span = blockIdx.y;
    const int z0 = blockIdx.z * rp.zgroup;
    const int z1 = min(z0 + rp.zgroup, rp.nz);
    const float2* fz = fzT + (size_t)f * M;
    const float2* fzh = fzT + (size_t)h * M;
    for (int z = z0; z < z1; ++z) {
        float2* col = rp.S + (((size_t)span * rp.nz + z) * h + f) * (size_t)M;
        const int2 ir = rp.in_rows ? rp.in_rows[span] : make_int2(0, M);
        stage_in(
            M, [&](int i) { return (i >= ir.x && i < ir.y) ? col[i] : make_float2(0.f, 0.f); },
            [&](int i, float2 v) { smc[lidx(i)] = v; });
        __syncthreads();
        fft_lds<EPT>(smc, M, 1, planM, twM, -1);
        if (f == 0) {
            for (int k = threadIdx.x; k <= M / 2; k += blockDim.x) {
                int km = k == 0 ? 0 : M - k;
                float2 A = smc[lidx(k)], Bc = cconj(smc[lidx(km)]);
                float2 D = make_float2(0.5f * (A.x + Bc.x), 0.5f * (A.y + Bc.y));
                float2 E = make_float2(0.5f * (A.x - Bc.x), 0.5f * (A.y - Bc.y));
                float2 Q = make_float2(E.y, -E.x);  // E / i
                float2 yk = cadd(cmul(D, fzval(fz, k, rp.conj)), rot90(cmul(Q, fzval(fzh, k, rp.conj)), 1));
                float2 ykm = cadd(cmul(cconj(D), fzval(fz, km, rp.conj)), rot90(cmul(cconj(Q), fzval(fzh, km, rp.conj)), 1));
                smc[lidx(k)] = yk;
                if (km != k) smc[lidx(km)] = ykm;
            }
        } else {
            for (int i = threadIdx.x; i < M; i += blockDim.x) smc[lidx(i)] = cmul(smc[lidx(i)], fzval(fz, i, rp.conj));
        }
        __syncthreads();
        fft_lds<EPT>(smc, M, 1, planM, twM, +1);
        const unsigned char* om = rp.out_rows ? rp.out_rows + (size_t)span * M : nullptr;
        for (int i = threadIdx.x; i < M; i += blockDim.x)
            if (!om || om[i]) col[i] = smc[lidx(i)];
        __syncthreads();
    }
}

// w^q (q = 1..15) applied to v[q] from table reads of w, w^2, w^4, w^8 and at
// most three products each
__device__ __forceinline__ void twiddle16(float2* v, const float2* __restrict__ tw, int base, int dir) {
    float2 w1 = twiddle(tw, base, dir), w2 = twiddle(tw, 2 * base, dir);
    float2 w4 = twiddle(tw, 4 * base, dir), w8 = twiddle(tw, 8 * base, dir);
    float2 w3 = cmul(w1, w2);
    v[1] = cmul(v[1], w1);
    v[2] = cmul(v[2], w2);
    v[3] = cmul(v[3], w3);
    v[4] = cmul(v[4], w4);
    v[5] = cmul(v[5], cmul(w1, w4));
    v[6] = cmul(v[6], cmul(w2, w4));
    v[7] = cmul(v[7], cmul(w3, w4));
    v[8] = cmul(v[8], w8);
    v[9] = cmul(v[9], cmul(w1, w8));
    v[10] = cmul(v[10], cmul(w2, w8));
    v[11] = cmul(v[11], cmul(w3, w8));
    float2 w12 = cmul(w4, w8);
    v[12] = cmul(v[12], w12);
    v[13] = cmul(v[13], cmul(w1, w12));
    v[14] = cmul(v[14], cmul(w2, w12));
    v[15] = cmul(v[15], cmul(w3, w12));
}

// Column pass for power-of-two M = 16 * MID * 16 (MID in 1, 2, 4, 8, 16):
// M/16 threads per column hold one radix-16 butterfly each.  Stage 0 reads
// HBM directly, the last forward stage leaves thread i holding outputs
// i + q M/16 -- exactly the inputs of the first inverse stage -- so the
// kernel-spectrum multiply and the inverse start in registers; 4 LDS
// exchanges in total.  C columns share a 256-thread block for small M.
template <int MID>
__device__ __forceinline__ void col_mid_stage(float2* buf, int M, int dir, const float2* __restrict__ tw, int li,
                                              bool worker) {
    // radix-MID stage, Ns = 16; thread li handles 16/MID butterflies
    constexpr int NB = 16 / MID;
    const int nbf = M / MID;
    float2 v[NB][MID];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        int j = li * NB + b;
#pragma unroll
        for (int q = 0; q < MID; ++q) v[b][q] = buf[lidx(j + q * nbf)];
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        int j = li * NB + b;
        int k = j & 15;
        const int tstep = M / (16 * MID);
        if (MID == 16) {
            twiddle16(v[b], tw, k * tstep, dir);   // powers formed as they are applied
        } else if (MID > 1) {
            float2 w[MID > 1 ? MID : 2];
            twiddle_pows<MID>(tw, k * tstep, dir, w);
#pragma unroll
            for (int q = 1; q < MID; ++q) v[b][q] = cmul(v[b][q], w[q]);
        }
        if (MID == 16) dft16(v[b], dir);
        else if (MID == 8) dft_small<8>(v[b], dir);
        else if (MID == 4) dft_small<4>(v[b], dir);
        else if (MID == 2) dft_small<2>(v[b], dir);
        int dst = (j - k) * MID + k;
        if (worker) {
#pragma unroll
            for (int q = 0; q < MID; ++q) buf[lidx(dst + q * 16)] = v[b][q];
        }
    }
    __syncthreads();
}


template <int MID>
__global__ void __launch_bounds__(256, 2) k_cols16(RowPass rp, const float2* __restrict__ twM,
                                                   const float2* __restrict__ fzT) {
    extern __shared__ float2 smc[];
    // one column per block; block-uniform base pointers with 32-bit lane
    // offsets keep the 16 addresses out of the vector registers
    constexpr int M = 256 * MID;            // == rp.Nrho (checked on launch)
    constexpr int T16 = M / 16;             // working threads per column
    const int h = rp.h;
    const int li = min((int)threadIdx.x, T16 - 1);
    const bool worker = (int)threadIdx.x < T16;
    // logical order: slice groups, then spans of one column are neighbours
    // (they read the same kernel-spectrum column); column 0 runs on k_cols
    const int nzg = (rp.nz + rp.zgroup - 1) / rp.zgroup;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int zg = lb % nzg, t = lb / nzg;
    const int span = t % 3, f = 1 + t / 3;
    float2* buf = smc;
    const float2* fz = fzT + (size_t)f * M;
    const int z0 = zg * rp.zgroup;
    const int z1 = min(z0 + rp.zgroup, rp.nz);
    constexpr int tstep3 = M / (16 * MID * 16);  // last stage: Ns = 16 MID, radix 16 (= 1)
    const int2 ir = rp.in_rows ? rp.in_rows[span] : make_int2(0, M);
    unsigned smask = 0xffffu;   // stored rows among li + q T16, read once
    if (rp.out_rows) {
        const unsigned char* om = rp.out_rows + (size_t)span * M;
        smask = 0;
#pragma unroll
        for (int q = 0; q < 16; ++q) smask |= om[li + q * T16] ? (1u << q) : 0u;
    }
    for (int z = z0; z < z1; ++z) {
        float2* col = rp.S + (((size_t)span * rp.nz + z) * h + f) * (size_t)M;
        float2 v[16];
        // ---- forward stage 0: radix 16, Ns = 1, straight from HBM (rows
        // without data are zeros)
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int i = li + q * T16;
            v[q] = (i >= ir.x && i < ir.y) ? col[i] : make_float2(0.f, 0.f);
        }
        dft16(v, -1);
        if (worker) {
#pragma unroll
            for (int q = 0; q < 16; ++q) buf[lidx(li * 16 + q)] = v[q];
        }
        __syncthreads();
        if (MID > 1) col_mid_stage<MID>(buf, M, -1, twM, li, worker);
        // ---- forward last stage: radix 16, Ns = M/16, k = li
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = buf[lidx(li + q * T16)];
        twiddle16(v, twM, li * tstep3, -1);
        dft16(v, -1);
        // v[q] = spectrum at kappa = li + q M/16: kernel multiply in registers
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = cmul(v[q], fzval(fz, li + q * T16, rp.conj));
        // ---- inverse stage 0 (radix 16, Ns = 1)
        dft16(v, +1);
        __syncthreads();
        if (worker) {
#pragma unroll
            for (int q = 0; q < 16; ++q) buf[lidx(li * 16 + q)] = v[q];
        }
        __syncthreads();
        if (MID > 1) col_mid_stage<MID>(buf, M, +1, twM, li, worker);
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = buf[lidx(li + q * T16)];
        twiddle16(v, twM, li * tstep3, +1);
        dft16(v, +1);
        if (worker) {
#pragma unroll
            for (int q = 0; q < 16; ++q)
                if (smask & (1u << q)) col[li + q * T16] = v[q];
        }
        __syncthreads();
    }
}

// Block barrier that waits only for this wave's LDS traffic (lgkmcnt), not
// for its outstanding global loads: the load-to-LDS stream of the next column
// keeps running across it.  Compiler-level ordering from the wave-scope fences.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_s_waitcnt(0xc07f);   // lgkmcnt(0)
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// w^q (q = 1..15) applied to v[q] from w, w^2, w^4, w^8 held in registers
// (conjugated for the inverse)
__device__ __forceinline__ void twiddle16_w(float2* v, const float2* w, int dir) {
    const bool cj = dir > 0;
    float2 w1 = cj ? cconj(w[0]) : w[0], w2 = cj ? cconj(w[1]) : w[1];
    float2 w4 = cj ? cconj(w[2]) : w[2], w8 = cj ? cconj(w[3]) : w[3];
    float2 w3 = cmul(w1, w2);
    v[1] = cmul(v[1], w1);
    v[2] = cmul(v[2], w2);
    v[3] = cmul(v[3], w3);
    v[4] = cmul(v[4], w4);
    v[5] = cmul(v[5], cmul(w1, w4));
    v[6] = cmul(v[6], cmul(w2, w4));
    v[7] = cmul(v[7], cmul(w3, w4));
    v[8] = cmul(v[8], w8);
    v[9] = cmul(v[9], cmul(w1, w8));
    v[10] = cmul(v[10], cmul(w2, w8));
    v[11] = cmul(v[11], cmul(w3, w8));
    float2 w12 = cmul(w4, w8);
    v[12] = cmul(v[12], w12);
    v[13] = cmul(v[13], cmul(w1, w12));
    v[14] = cmul(v[14], cmul(w2, w12));
    v[15] = cmul(v[15], cmul(w3, w12));
}

// Column pass for M = 4096 (16 x 16 x 16) with the next slice's column
// streamed from HBM into LDS (load-to-LDS, no registers) while the current one
// is transformed.  The twiddles and the kernel-spectrum column are loaded once
// per block, so no other global load shares the vector-memory counter with the
// stream and every barrier inside a slice waits for LDS traffic only.
__global__ void __launch_bounds__(256, 2) k_cols4096(RowPass rp, const float2* __restrict__ twM,
                                                     const float2* __restrict__ fzT) {
    constexpr int M = 4096, T16 = 256;
    __shared__ float2 buf[lds_slots(M)];
    __shared__ __align__(16) float2 pre[M];
    const int h = rp.h;
    const int li = threadIdx.x;
    const int nzg = (rp.nz + rp.zgroup - 1) / rp.zgroup;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int zg = lb % nzg, t = lb / nzg;
    const int span = t % 3, f = 1 + t / 3;
    const int z0 = zg * rp.zgroup;
    const int z1 = min(z0 + rp.zgroup, rp.nz);
    const int wv = li >> 6, ln = li & 63;
    // 128-row chunks of the column that carry data (a multiple of 128 bounds
    // the rows with data); the others stay zero in `pre`
    const int2 ir = rp.in_rows ? rp.in_rows[span] : make_int2(0, M);
    const int c_lo = ir.x >> 7, c_hi = (ir.y + 127) >> 7;
    // rows of this thread's 16 outputs (li + q 256) that are stored: read once,
    // before the stream starts, so that no other global load joins it
    unsigned smask = 0xffffu;
    if (rp.out_rows) {
        const unsigned char* om = rp.out_rows + (size_t)span * M;
        smask = 0;
#pragma unroll
        for (int q = 0; q < 16; ++q) smask |= om[li + q * T16] ? (1u << q) : 0u;
    }
    for (int i = li; i < M; i += T16)
        if ((i >> 7) < c_lo || (i >> 7) >= c_hi) pre[i] = make_float2(0.f, 0.f);
    auto prefetch = [&](int z) {
        const float2* col = rp.S + (((size_t)span * rp.nz + z) * h + f) * (size_t)M;
#pragma unroll
        for (int c = 0; c < M / 128 / 4; ++c) {
            const int cc = wv + 4 * c;
            if (cc >= c_lo && cc < c_hi)
                __builtin_amdgcn_global_load_lds(col + cc * 128 + ln * 2, pre + cc * 128, 16, 0, kCpolNT);
        }
    };
    prefetch(z0);
    // per-block constants: middle-stage twiddles (k = li & 15, step 16),
    // last-stage twiddles (k = li, step 1), the kernel-spectrum column
    float2 wm[4], wl[4], kz[16];
    {
        const int km = (li & 15) * 16;
        wm[0] = twM[km]; wm[1] = twM[2 * km]; wm[2] = twM[4 * km]; wm[3] = twM[8 * km];
        wl[0] = twM[li]; wl[1] = twM[2 * li]; wl[2] = twM[4 * li]; wl[3] = twM[8 * li];
        const float2* fz = fzT + (size_t)f * M;
#pragma unroll
        for (int q = 0; q < 16; ++q) kz[q] = fzval(fz, li + q * T16, rp.conj);
    }
    for (int z = z0; z < z1; ++z) {
        float2* col = rp.S + (((size_t)span * rp.nz + z) * h + f) * (size_t)M;
        float2 v[16];
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();   // column z is in pre
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = pre[li + q * T16];
        dft16(v, -1);
#pragma unroll
        for (int q = 0; q < 16; ++q) buf[lidx(li * 16 + q)] = v[q];
        lds_barrier();   // pre has been read by every wave
        if (z + 1 < z1) prefetch(z + 1);
        // middle stage (radix 16, Ns = 16)
        {
            const int j = li, k = j & 15;
#pragma unroll
            for (int q = 0; q < 16; ++q) v[q] = buf[lidx(j + q * 256)];
            lds_barrier();
            twiddle16_w(v, wm, -1);
            dft16(v, -1);
            const int dst = (j - k) * 16 + k;
#pragma unroll
            for (int q = 0; q < 16; ++q) buf[lidx(dst + q * 16)] = v[q];
            lds_barrier();
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = buf[lidx(li + q * T16)];
        twiddle16_w(v, wl, -1);
        dft16(v, -1);
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = cmul(v[q], kz[q]);
        dft16(v, +1);
        lds_barrier();
#pragma unroll
        for (int q = 0; q < 16; ++q) buf[lidx(li * 16 + q)] = v[q];
        lds_barrier();
        {
            const int j = li, k = j & 15;
#pragma unroll
            for (int q = 0; q < 16; ++q) v[q] = buf[lidx(j + q * 256)];
            lds_barrier();
            twiddle16_w(v, wm, +1);
            dft16(v, +1);
            const int dst = (j - k) * 16 + k;
#pragma unroll
            for (int q = 0; q < 16; ++q) buf[lidx(dst + q * 16)] = v[q];
            lds_barrier();
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = buf[lidx(li + q * T16)];
        twiddle16_w(v, wl, +1);
        dft16(v, +1);
#pragma unroll
        for (int q = 0; q < 16; ++q)
            if (smask & (1u << q)) st_at_nt(col, (unsigned)(li + q * T16) * 8u, v[q]);
        lds_barrier();
    }
}

// Barrier over the lanes that share a row.  When a row fits in one wave
// (h <= 1024) its lanes execute together and a wave's LDS operations complete
// in issue order, so a wave-scope fence pair around a compiler-level wave
// barrier orders the exchange; wider rows need the block barrier.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
template <bool WAVE>
__device__ __forceinline__ void row_sync() {
    if constexpr (WAVE) wave_sync();
    else __syncthreads();
}

// r2c_post / c2r_pre of one row by the T16 lanes that own it, unrolled:
// lane li owns the pairs (k, H - k) for k = li + i T16 < H/2 and lane 0 also
// k = H/2.  Every slot is read and written by its owner only, so no barrier
// separates the reads from the writes; the twiddle loads are all issued
// before the first use.
template <int H, int T16>
__device__ __forceinline__ void r2c_post_row(float2* row, const float2* __restrict__ twN, int li) {
    constexpr int NI = H / (2 * T16);
    float2 W[NI], A[NI], B[NI];
#pragma unroll
    for (int i = 0; i < NI; ++i) {
        const int k = li + i * T16;
        W[i] = twN[k];
        A[i] = row[lidx(k)];
        B[i] = row[lidx(k == 0 ? 0 : H - k)];
    }
    float2 Am = make_float2(0.f, 0.f), Wm = make_float2(0.f, 0.f);
    if (li == 0) { Am = row[lidx(H / 2)]; Wm = twN[H / 2]; }
#pragma unroll
    for (int i = 0; i < NI; ++i) {
        const int k = li + i * T16;
        if (k == 0) {
            row[0] = make_float2(A[i].x + A[i].y, A[i].x - A[i].y);
        } else {
            float2 Bc = cconj(B[i]);
            float2 E = make_float2(0.5f * (A[i].x + Bc.x), 0.5f * (A[i].y + Bc.y));
            float2 D = make_float2(0.5f * (A[i].x - Bc.x), 0.5f * (A[i].y - Bc.y));
            float2 WO = cmul(W[i], make_float2(D.y, -D.x));
            row[lidx(k)] = cadd(E, WO);
            row[lidx(H - k)] = cconj(csub(E, WO));
        }
    }
    if (li == 0) {
        // k = H/2 pairs with itself: E = Re A, D = i Im A
        float2 WO = cmul(Wm, make_float2(Am.y, 0.f));
        row[lidx(H / 2)] = cadd(make_float2(Am.x, 0.f), WO);
    }
}
template <int H, int T16>
__device__ __forceinline__ void c2r_pre_row(float2* row, const float2* __restrict__ twN, int li) {
    constexpr int NI = H / (2 * T16);
    float2 W[NI], P[NI], Q[NI];
#pragma unroll
    for (int i = 0; i < NI; ++i) {
        const int k = li + i * T16;
        W[i] = twN[k];
        P[i] = row[lidx(k)];
        Q[i] = row[lidx(k == 0 ? 0 : H - k)];
    }
    float2 Pm = make_float2(0.f, 0.f), Wm = make_float2(0.f, 0.f);
    if (li == 0) { Pm = row[lidx(H / 2)]; Wm = twN[H / 2]; }
#pragma unroll
    for (int i = 0; i < NI; ++i) {
        const int k = li + i * T16;
        if (k == 0) {
            row[0] = make_float2(P[i].x + P[i].y, P[i].x - P[i].y);
        } else {
            float2 S = cadd(P[i], cconj(Q[i]));
            float2 D = csub(P[i], cconj(Q[i]));
            row[lidx(k)] = cadd(S, cmul(make_float2(W[i].y, W[i].x), D));
            row[lidx(H - k)] = cadd(cconj(S), cmul(make_float2(W[i].y, -W[i].x), cconj(make_float2(-D.x, -D.y))));
        }
    }
    if (li == 0) {
        float2 S = cadd(Pm, cconj(Pm)), D = csub(Pm, cconj(Pm));
        row[lidx(H / 2)] = cadd(S, cmul(make_float2(Wm.y, Wm.x), D));
    }
}

// Row passes for power-of-two h = 16 * MID * 16: T16 = h/16 threads per row,
// R rows per block, one radix-16 butterfly per thread per stage.
template <int MID, bool WAVE>
__device__ __forceinline__ void row_mid_stage(float2* rowbuf, int dir, const float2* __restrict__ tw, int li) {
    // radix-MID stage (Ns = 16) of one row held at rowbuf (lidx-padded)
    constexpr int H = 256 * MID;
    constexpr int NB = 16 / MID;
    constexpr int nbf = H / MID;
    float2 v[NB][MID];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        int j = li * NB + b;
#pragma unroll
        for (int q = 0; q < MID; ++q) v[b][q] = rowbuf[lidx(j + q * nbf)];
    }
    row_sync<WAVE>();
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        int j = li * NB + b;
        int k = j & 15;
        if (MID == 16) {
            twiddle16(v[b], tw, k * (H / (16 * MID)), dir);   // powers formed as they are applied
        } else if (MID > 1) {
            float2 w[MID > 1 ? MID : 2];
            twiddle_pows<MID>(tw, k * (H / (16 * MID)), dir, w);
#pragma unroll
            for (int q = 1; q < MID; ++q) v[b][q] = cmul(v[b][q], w[q]);
        }
        if (MID == 16) dft16(v[b], dir);
        else if (MID == 8) dft_small<8>(v[b], dir);
        else if (MID == 4) dft_small<4>(v[b], dir);
        else if (MID == 2) dft_small<2>(v[b], dir);
        int dst = (j - k) * MID + k;
#pragma unroll
        for (int q = 0; q < MID; ++q) rowbuf[lidx(dst + q * 16)] = v[b][q];
    }
    row_sync<WAVE>();
}

template <int MID>
__global__ void __launch_bounds__(1024) k_rows16_fwd(RowPass rp, const float* __restrict__ F,
                                                     const int2* __restrict__ range) {
    extern __shared__ float2 smc[];
    constexpr int H = 256 * MID, T16 = 16 * MID;
    constexpr bool WAVE = T16 <= 64;   // a row's lanes lie in one wave
    const int Nt = rp.Ntheta, R = rp.R, Nr = rp.Nrho;
    // row stride in LDS slots: the padded row plus 32/R slots.  The transposed
    // accesses rr * RS + lidx(f) of a half-wave cover R rows x 32/R columns,
    // which then land on 32 distinct 8-byte bank pairs
    const int RS = H + H / 16 + 32 / R;
    const int lgR = __builtin_ctz((unsigned)R);   // R is a power of two
    const int r = threadIdx.x / T16, li = threadIdx.x - r * T16;
    // logical order: neighbouring row blocks (which share the 128-byte lines
    // of the transposed spectra) run behind the same L2
    const int nrb = Nr / R, nzg = (rp.nz + rp.zgroup - 1) / rp.zgroup;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int row0 = (lb % nrb) * R, span = lb / (nrb * nzg);
    const int z0 = ((lb / nrb) % nzg) * rp.zgroup;
    const int z1 = min(z0 + rp.zgroup, rp.nz);
    if (rp.in_rows) {
        // no data in these rows: the column pass reads them as zeros
        const int2 ir = rp.in_rows[span];
        if (row0 + R <= ir.x || row0 >= ir.y) return;
    }
    const int2 lr = range ? range[(size_t)span * Nr + row0 + r] : make_int2(0, Nt);
    float2* rowbuf = smc + r * RS;
    // Up to H = 2048 the block's twiddles (twh[0, H/2), twN[0, H/2]) fit in
    // LDS after the rows: loaded once, they spare every slice three waits on
    // L2 after its barriers.
    constexpr bool TWS = MID <= 8;
    const float2* twh = rp.twh;
    const float2* twN = rp.twN;
    if constexpr (TWS) {
        float2* ts = smc + R * RS;
        for (int i = threadIdx.x; i <= H; i += blockDim.x) ts[i] = i < H / 2 ? rp.twh[i] : rp.twN[i - H / 2];
        __syncthreads();
        twh = ts;
        twN = ts + H / 2;
    }
    for (int z = z0; z < z1; ++z) {
        // all 16 loads in flight at once: unconditional, at a row base plus a
        // 32-bit lane offset clamped into the sampled range (the clamped pairs
        // lie in lines that are loaded anyway); pairs outside it are zeroed
        // The index math is made opaque per slice so that the compiler does not
        // keep 16 hoisted addresses and masks live across the loop: at the
        // 128-VGPR budget of a 1024-lane block it would then issue the loads
        // one at a time.
        const float* Fr = F + ((((size_t)span * rp.nz + z) * Nr + row0 + r) * (size_t)Nt);
        int l2 = 2 * li, rx = lr.x, ry = lr.y;
        asm volatile("" : "+v"(l2), "+v"(rx), "+v"(ry));
        const int tlo = min(max(rx & ~1, 0), Nt - 2), thi = max(tlo, min((ry - 1) & ~1, Nt - 2));
        float2 v[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int th = l2 + 2 * q * T16;
            v[q] = ld_at(reinterpret_cast<const float2*>(Fr), (unsigned)min(max(th, tlo), thi) * 4u);
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int th = l2 + 2 * q * T16;
            v[q].x = (th >= rx && th < ry) ? v[q].x : 0.f;
            v[q].y = (th + 1 >= rx && th + 1 < ry) ? v[q].y : 0.f;
        }
        dft16(v, -1);
#pragma unroll
        for (int q = 0; q < 16; ++q) rowbuf[lidx(li * 16 + q)] = v[q];
        row_sync<WAVE>();
        if (MID > 1) row_mid_stage<MID, WAVE>(rowbuf, -1, twh, li);
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = rowbuf[lidx(li + q * T16)];
        twiddle16(v, twh, li, -1);
        dft16(v, -1);
        row_sync<WAVE>();
#pragma unroll
        for (int q = 0; q < 16; ++q) rowbuf[lidx(li + q * T16)] = v[q];
        row_sync<WAVE>();
        r2c_post_row<H, T16>(rowbuf, twN, li);
        lds_barrier();   // the transposed store reads every row
        {
            // thread (rr, f0) = (tid mod R, tid / R) stores frequencies f0 + i T16
            // (i < 16) of row rr: LDS slots at fixed steps, global addresses a
            // uniform base per i plus one lane offset
            const int rr = threadIdx.x & (R - 1), f0 = threadIdx.x >> lgR;
            const float2* src = smc + rr * RS + lidx(f0);
            float2* out = rp.S + ((size_t)span * rp.nz + z) * H * (size_t)Nr + row0;
            const unsigned lane = (unsigned)(f0 * Nr + rr) * 8u;
            float2 w[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) w[i] = src[i * (T16 + T16 / 16)];
#pragma unroll
            for (int i = 0; i < 16; ++i) st_at(out + (size_t)(i * T16) * Nr, lane, w[i]);
        }
        // LDS reuse only: the stores of S drain while the next slice runs
        lds_barrier();
    }
}

template <int MID>
__global__ void __launch_bounds__(1024) k_rows16_inv(RowPass rp, float* __restrict__ G,
                                                     const int2* __restrict__ wrange) {
    extern __shared__ float2 smc[];
    constexpr int H = 256 * MID, T16 = 16 * MID;
    constexpr bool WAVE = T16 <= 64;
    const int Nt = rp.Ntheta, R = rp.R, Nr = rp.Nrho;
    const int RS = H + H / 16 + 32 / R;
    const int lgR = __builtin_ctz((unsigned)R);   // R is a power of two
    const int r = threadIdx.x / T16, li = threadIdx.x - r * T16;
    const int nrb = Nr / R, nzg = (rp.nz + rp.zgroup - 1) / rp.zgroup;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int row0 = (lb % nrb) * R, span = lb / (nrb * nzg);
    const int z0 = ((lb / nrb) % nzg) * rp.zgroup;
    const int z1 = min(z0 + rp.zgroup, rp.nz);
    // theta columns of this row that the output gather reads (null: all), as
    // a range of float2 slots
    const int2 wr = wrange ? wrange[(size_t)span * Nr + row0 + r] : make_int2(0, Nt);
    const int e0 = max(wr.x, 0) >> 1, e1 = min((wr.y + 1) >> 1, H);
    // no row of the block is read by the output gather (the column pass did
    // not store these rows either)
    if (rp.out_rows && !__syncthreads_or(e0 < e1)) return;
    float2* rowbuf = smc + r * RS;
    // (k_rows16_fwd stages its twiddles in LDS; the same staging measured
    // 4 % slower here, so this pass reads them from the tables)
    constexpr bool TWS = false;
    const float2* twh = rp.twh;
    const float2* twN = rp.twN;
    if constexpr (TWS) {
        float2* ts = smc + R * RS;
        for (int i = threadIdx.x; i <= H; i += blockDim.x) ts[i] = i < H / 2 ? rp.twh[i] : rp.twN[i - H / 2];
        __syncthreads();
        twh = ts;
        twN = ts + H / 2;
    }
    for (int z = z0; z < z1; ++z) {
        {
            // transposed load, the mirror of k_rows16_fwd's store
            const int rr = threadIdx.x & (R - 1), f0 = threadIdx.x >> lgR;
            float2* dst = smc + rr * RS + lidx(f0);
            const float2* in = rp.S + ((size_t)span * rp.nz + z) * H * (size_t)Nr + row0;
            const unsigned lane = (unsigned)(f0 * Nr + rr) * 8u;
            float2 w[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) w[i] = ld_at_nt(in + (size_t)(i * T16) * Nr, lane);
#pragma unroll
            for (int i = 0; i < 16; ++i) dst[i * (T16 + T16 / 16)] = w[i];
        }
        lds_barrier();   // every row was written by the whole block
        c2r_pre_row<H, T16>(rowbuf, twN, li);
        row_sync<WAVE>();
        float2 v[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = rowbuf[lidx(li + q * T16)];
        dft16(v, +1);
        row_sync<WAVE>();
#pragma unroll
        for (int q = 0; q < 16; ++q) rowbuf[lidx(li * 16 + q)] = v[q];
        row_sync<WAVE>();
        if (MID > 1) row_mid_stage<MID, WAVE>(rowbuf, +1, twh, li);
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = rowbuf[lidx(li + q * T16)];
        twiddle16(v, twh, li, +1);
        dft16(v, +1);
        // back through the row's LDS slots, then only the columns read later
        row_sync<WAVE>();
#pragma unroll
        for (int q = 0; q < 16; ++q) rowbuf[lidx(li + q * T16)] = v[q];
        row_sync<WAVE>();
        float2* out = reinterpret_cast<float2*>(G + ((((size_t)span * rp.nz + z) * Nr + row0 + r) * (size_t)Nt));
        for (int e = e0 + li; e < e1; e += T16) st_at_nt(out, (unsigned)e * 8u, rowbuf[lidx(e)]);
        lds_barrier();   // the next slice's staging overwrites every row (stores drain meanwhile)
    }
}

// Length-32 DFT in registers, natural order in and out: one radix-2 step
// over two register radix-16 DFTs (even and odd elements), twiddles
// W32^k = exp(dir 2 pi i k / 32) as constants.
__device__ __forceinline__ void dft32(float2* v, int dir) {
    // cos and sin of pi k / 16
    constexpr float C[16] = {1.0f, 0.98078528040323044913f, 0.92387953251128675613f, 0.83146961230254523708f,
                             0.70710678118654752440f, 0.55557023301960222474f, 0.38268343236508977173f,
                             0.19509032201612826785f, 0.0f, -0.19509032201612826785f, -0.38268343236508977173f,
                             -0.55557023301960222474f, -0.70710678118654752440f, -0.83146961230254523708f,
                             -0.92387953251128675613f, -0.98078528040323044913f};
    constexpr float S[16] = {0.0f, 0.19509032201612826785f, 0.38268343236508977173f, 0.55557023301960222474f,
                             0.70710678118654752440f, 0.83146961230254523708f, 0.92387953251128675613f,
                             0.98078528040323044913f, 1.0f, 0.98078528040323044913f, 0.92387953251128675613f,
                             0.83146961230254523708f, 0.70710678118654752440f, 0.55557023301960222474f,
                             0.38268343236508977173f, 0.19509032201612826785f};
    float2 e[16], o[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        e[k] = v[2 * k];
        o[k] = v[2 * k + 1];
    }
    dft16(e, dir);
    dft16(o, dir);
    const float d = (float)dir;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        float2 t;
        if (k == 0) t = o[0];
        else if (k == 8) t = rot90(o[8], dir);
        else t = cmul(o[k], make_float2(C[k], d * S[k]));
        v[k] = cadd(e[k], t);
        v[k + 16] = csub(e[k], t);
    }
}

// Inverse row pass at h = 1024 (Ntheta = 2048) with 32 lanes per row: a lane
// holds 32 elements, so the length-1024 FFT is one register DFT32 over
// elements 32 apart, the twiddles W1024^(l m), one LDS exchange within the
// row and a second register DFT32 (1024 = 32 x 32).  16 rows per 512-lane
// block (the transposed half-spectra are read in 128-byte segments); at 512
// lanes a lane has 256 VGPRs, which hold the next slice's 32 loads while the
// current slice is transformed and stored.  S is read and G written with the
// non-temporal hint.  (The same scheme for the forward row pass measured
// slower than k_rows16_fwd, with or without the prefetch.)
// LDS: 16 rows at stride 1090 slots (the layout of k_rows16_*), then the
// twiddles W1024^(l m) at [m * 32 + l] (a fixed m is read conflict-free by
// the 32 lanes l) and twN[0, 512].
constexpr int kRows32RS = 1024 + 1024 / 16 + 2;
constexpr size_t kRows32Lds = (size_t)(16 * kRows32RS + 1024 + 513) * sizeof(float2);

__device__ __forceinline__ void rows32_tables(float2* tw2, float2* twn, const RowPass& rp) {
    for (int i = threadIdx.x; i < 1024; i += 512) tw2[i] = rp.twh[(i & 31) * (i >> 5)];
    for (int i = threadIdx.x; i <= 512; i += 512) twn[i] = rp.twN[i];
}

// Length-1024 FFT of the row owned by lanes li = 0..31 of one wave half,
// v[q] = x[li + 32 q] on entry, v[k] = X[li + 32 k] on exit.  The exchange
// goes through the row's slots m * 33 + l (m, l < 32; the odd stride keeps
// both the writes, fixed m, and the reads, fixed l, free of bank
// conflicts), which must be free on entry and are free again on exit.
__device__ __forceinline__ void rows32_fft(float2* v, float2* row, const float2* tw2, int li, int dir) {
    dft32(v, dir);
#pragma unroll
    for (int m = 1; m < 32; ++m) {
        const float2 w = tw2[m * 32 + li];
        v[m] = cmul(v[m], dir < 0 ? w : cconj(w));
    }
#pragma unroll
    for (int m = 0; m < 32; ++m) row[m * 33 + li] = v[m];
    wave_sync();
#pragma unroll
    for (int l = 0; l < 32; ++l) v[l] = row[li * 33 + l];
    dft32(v, dir);
    wave_sync();   // the exchange slots are read before the caller reuses them
}

__global__ void __launch_bounds__(512) k_rows32_inv(RowPass rp, float* __restrict__ G,
                                                    const int2* __restrict__ wrange) {
    extern __shared__ float2 smc[];
    constexpr int H = 1024, T = 32, R = 16, RS = kRows32RS;
    const int Nt = rp.Ntheta, Nr = rp.Nrho;
    const int r = threadIdx.x >> 5, li = threadIdx.x & 31;
    const int nrb = Nr / R, nzg = (rp.nz + rp.zgroup - 1) / rp.zgroup;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int row0 = (lb % nrb) * R, span = lb / (nrb * nzg);
    const int z0 = ((lb / nrb) % nzg) * rp.zgroup;
    const int z1 = min(z0 + rp.zgroup, rp.nz);
    const int2 wr = wrange ? wrange[(size_t)span * Nr + row0 + r] : make_int2(0, Nt);
    const int e0 = max(wr.x, 0) >> 1, e1 = min((wr.y + 1) >> 1, H);
    if (rp.out_rows && !__syncthreads_or(e0 < e1)) return;
    float2* row = smc + r * RS;
    float2* tw2 = smc + R * RS;
    float2* twn = tw2 + 1024;
    rows32_tables(tw2, twn, rp);
    // transposed load: thread (rr, f0) = (tid mod 16, tid / 16) loads
    // frequencies f0 + 32 i of row rr
    const int rr = threadIdx.x & (R - 1), f0 = threadIdx.x >> 4;
    const unsigned lane = (unsigned)(f0 * Nr + rr) * 8u;
    float2* dst = smc + rr * RS + lidx(f0);
    const unsigned step = 32u * (unsigned)Nr * 8u;   // bytes between frequencies f0 + 32 i and f0 + 32 (i + 1)
    auto load = [&](int z, float2* w) {
        const float2* in = rp.S + ((size_t)span * rp.nz + z) * H * (size_t)Nr + row0;
        unsigned off = lane;
        asm volatile("" : "+v"(off));
#pragma unroll
        for (int i = 0; i < 32; ++i) w[i] = ld_at_nt(in, off + (unsigned)i * step);
    };
    float2 v[32], nx[32];
    load(z0, nx);
    for (int z = z0; z < z1; ++z) {
#pragma unroll
        for (int i = 0; i < 32; ++i) dst[34 * i] = nx[i];
        lds_barrier();   // every row was written by the whole block (and the tables, first time)
        // the lane index, opaque per slice: the LDS addresses derived from it
        // are recomputed here instead of being held across the loop next to
        // the prefetched loads
        int l = li;
        asm volatile("" : "+v"(l));
        c2r_pre_row<H, T>(row, twn, l);
        wave_sync();
#pragma unroll
        for (int q = 0; q < 32; ++q) v[q] = row[lidx(l + 32 * q)];
        wave_sync();   // the exchange slots overlap the row
        if (z + 1 < z1) load(z + 1, nx);   // in flight while this slice is transformed and stored
        rows32_fft(v, row, tw2, l, +1);
        // v[k] = x[li + 32 k]: only the columns the output gather reads
        float2* out = reinterpret_cast<float2*>(G + ((((size_t)span * rp.nz + z) * Nr + row0 + r) * (size_t)Nt));
#pragma unroll
        for (int k = 0; k < 32; ++k) {
            const int e = l + 32 * k;
            if (e >= e0 && e < e1) st_at_nt(out, (unsigned)e * 8u, v[k]);
        }
        lds_barrier();   // the next slice's staging overwrites every row
    }
}

// Inverse row pass fused with the log-polar -> polar gather of the forward
// operator.  A block owns a band of `nbase` = R - 3 base rows of one span:
// it inverse-transforms the R rows [c0, c0 + R) (c0 = band * nbase - 1, mod
// Nrho), which hold every cubic tap of the polar samples whose floor(rho)
// lies in the band, keeps the real rows in LDS and evaluates those samples
// from there, so the convolved grid G never goes through HBM.  At fixed angle
// rho' grows with the detector coordinate, so a band's samples are runs of
// consecutive detector cells per angle (listed on the host in that order).
template <int MID>
__global__ void __launch_bounds__(1024) k_rows16_inv_polar(RowPass rp, PolarOut po) {
    extern __shared__ float2 smc[];
    constexpr int H = 256 * MID, T16 = 16 * MID;
    constexpr bool WAVE = T16 <= 64;
    const int Nt = rp.Ntheta, R = rp.R, Nr = rp.Nrho;
    const int RS = H + H / 16 + 32 / R;
    const int lgR = __builtin_ctz((unsigned)R);
    const int nbase = R - 3;
    const int r = threadIdx.x / T16, li = threadIdx.x - r * T16;
    const int nzg = (rp.nz + rp.zgroup - 1) / rp.zgroup;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int band = lb % po.nband, span = lb / (po.nband * nzg);
    const int z0 = ((lb / po.nband) % nzg) * rp.zgroup;
    const int z1 = min(z0 + rp.zgroup, rp.nz);
    const int* bb = po.band_beg + (size_t)span * (po.nband + 1) + band;
    const int beg = bb[0], end = bb[1];
    if (beg >= end) return;   // no polar sample has its base row here
    const int c0 = band * nbase - 1;
    float2* rowbuf = smc + r * RS;
    const float* rowsf = reinterpret_cast<const float*>(smc);
    const int fstride = 2 * RS;   // floats per LDS row
    const float2* twh = rp.twh;
    const float2* twN = rp.twN;
    for (int z = z0; z < z1; ++z) {
        {
            // transposed load of rows (c0 + rr) mod Nrho
            const int rr = threadIdx.x & (R - 1), f0 = threadIdx.x >> lgR;
            float2* dst = smc + rr * RS + lidx(f0);
            const float2* in = rp.S + ((size_t)span * rp.nz + z) * H * (size_t)Nr;
            const unsigned lane = (unsigned)(f0 * Nr + ((c0 + rr) & (Nr - 1))) * 8u;
            float2 w[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) w[i] = ld_at(in + (size_t)(i * T16) * Nr, lane);
#pragma unroll
            for (int i = 0; i < 16; ++i) dst[i * (T16 + T16 / 16)] = w[i];
        }
        lds_barrier();
        c2r_pre_row<H, T16>(rowbuf, twN, li);
        row_sync<WAVE>();
        float2 v[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = rowbuf[lidx(li + q * T16)];
        dft16(v, +1);
        row_sync<WAVE>();
#pragma unroll
        for (int q = 0; q < 16; ++q) rowbuf[lidx(li * 16 + q)] = v[q];
        row_sync<WAVE>();
        if (MID > 1) row_mid_stage<MID, WAVE>(rowbuf, +1, twh, li);
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = rowbuf[lidx(li + q * T16)];
        twiddle16(v, twh, li, +1);
        dft16(v, +1);
        row_sync<WAVE>();
        // the row's real samples, unpadded: theta = 2e, 2e + 1 in slot e
#pragma unroll
        for (int q = 0; q < 16; ++q) rowbuf[li + q * T16] = v[q];
        lds_barrier();   // every row is complete
        float* Oz = po.out + (size_t)z * po.o_z;
        for (int i = beg + (int)threadIdx.x; i < end; i += blockDim.x) {
            const float2 c = po.coord[i];
            const int off = po.off[i];
            const float fx = floorf(c.x), fy = floorf(c.y);
            const float ax = c.x - fx, ay = c.y - fy;
            // local row of tap 0 (row floor(rho) - 1, mod Nrho): in [0, nbase)
            const int l0 = ((int)fy - band * nbase) & (Nr - 1);
            float val;
            if (po.cubic) {
                const int j0 = (int)fx - 1;
                const float* q0 = rowsf + l0 * fstride;
                float wx[4], wy[4];
                bspline_w6(ax, wx);
                bspline_w6(ay, wy);
                const int ja = j0 & (Nt - 1), jb = (j0 + 1) & (Nt - 1), jc = (j0 + 2) & (Nt - 1), jd = (j0 + 3) & (Nt - 1);
                val = 0.f;
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const float* qr = q0 + b * fstride;
                    val += wy[b] * (wx[0] * qr[ja] + wx[1] * qr[jb] + wx[2] * qr[jc] + wx[3] * qr[jd]);
                }
                val *= (1.0f / 36.0f);
            } else {
                const int ja = (int)fx & (Nt - 1), jb = ((int)fx + 1) & (Nt - 1);
                const float* q1 = rowsf + (l0 + 1) * fstride;
                const float* q2 = q1 + fstride;
                val = (1.f - ay) * ((1.f - ax) * q1[ja] + ax * q1[jb]) + ay * ((1.f - ax) * q2[ja] + ax * q2[jb]);
            }
            Oz[off] = val;
        }
        lds_barrier();   // the next slice's staging overwrites every row (stores drain meanwhile)
    }
}

// ---------------------------------------------------------------------------
int ept_for(long long points, int nt) {
    long long e = (points + nt - 1) / nt;
    if (e <= 4) return 4;
    if (e <= 8) return 8;
    if (e <= 16) return 16;
    if (e <= 32) return 32;
    return -1;
}

size_t rows_lds_bytes(int ldm, int R) { return (size_t)lds_slots(R * ldm) * sizeof(float2); }
size_t cols_lds_bytes(int M) { return (size_t)lds_slots(M) * sizeof(float2); }
size_t sino_lds_bytes(int Lh, int N) { return (size_t)lds_slots(Lh) * sizeof(float2) + (size_t)(2 * N) * sizeof(float); }

// dynamic LDS above 64 KiB (gfx950 has 160 KiB per workgroup) is opted into
// per kernel; a refusal surfaces as a launch error in check_launch
template <class K>
static void allow_lds(K k, size_t bytes) {
    if (bytes > 65536) (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

#define LP_DISPATCH_EPT(ept, KER, grid, block, lds, stream, ...)                                   \
    do {                                                                                          \
        switch (ept) {                                                                            \
            case 4: allow_lds(KER<4>, lds); hipLaunchKernelGGL(KER<4>, grid, block, lds, stream, __VA_ARGS__); break;   \
            case 8: allow_lds(KER<8>, lds); hipLaunchKernelGGL(KER<8>, grid, block, lds, stream, __VA_ARGS__); break;   \
            case 16: allow_lds(KER<16>, lds); hipLaunchKernelGGL(KER<16>, grid, block, lds, stream, __VA_ARGS__); break; \
            default: allow_lds(KER<32>, lds); hipLaunchKernelGGL(KER<32>, grid, block, lds, stream, __VA_ARGS__); break; \
        }                                                                                         \
    } while (0)

#define LP_DISPATCH_EPT2(ept, KER, B, grid, block, lds, stream, ...)                                                    \
    do {                                                                                                             \
        switch (ept) {                                                                                               \
            case 4: allow_lds(KER<4, B>, lds); hipLaunchKernelGGL((KER<4, B>), grid, block, lds, stream, __VA_ARGS__); break;   \
            case 8: allow_lds(KER<8, B>, lds); hipLaunchKernelGGL((KER<8, B>), grid, block, lds, stream, __VA_ARGS__); break;   \
            case 16: allow_lds(KER<16, B>, lds); hipLaunchKernelGGL((KER<16, B>), grid, block, lds, stream, __VA_ARGS__); break; \
            default: allow_lds(KER<32, B>, lds); hipLaunchKernelGGL((KER<32, B>), grid, block, lds, stream, __VA_ARGS__); break; \
        }                                                                                                            \
    } while (0)

// threads per FFT block: keep <= 16 points per thread (EPT 32 spills past
// the 128-VGPR budget of a 1024-thread-capable kernel)
static int threads_for(long long points) {
    int nt = 256;
    while (points > 16LL * nt && nt < 1024) nt *= 2;
    return nt;
}

void launch_image_rows(const float* f, float* fc, int nz, int N0, int N, bool cubic, hipStream_t s) {
    int nt = 256;
    hipLaunchKernelGGL(k_image_rows, dim3(N, nz), dim3(nt), 2 * N * sizeof(float), s, f, fc, N0, N,
                       N / 2 - N0 / 2, (int)cubic);
}

void launch_sino_rows(const float* R, float* Rc, int nz, const SinoPrep& sp, bool cubic, hipStream_t s) {
    if (!sp.H2) {
        hipLaunchKernelGGL((k_sino_rows<4, false>), dim3(cubic ? sp.P0 : sp.P, nz), dim3(256),
                           2 * (size_t)sp.N * sizeof(float), s, R, Rc, sp, (int)cubic);
        return;
    }
    // two kept angles per block, one complex FFT pair of length 2N
    if (sp.rtc_pair) {
        SinoPrep q = sp;
        int cb = (int)cubic;
        void* args[] = {&R, &Rc, &q, &cb};
        if (hipModuleLaunchKernel((hipFunction_t)q.rtc_pair, (sp.P0 + 1) / 2, nz, 1, q.rtc_pair_nt, 1, 1, 0, s, args,
                                  nullptr) == hipSuccess)
            return;
        (void)hipGetLastError();   // the generic kernel below does the pass
    }
    int nt = threads_for(2LL * sp.N);
    int ept = ept_for(2LL * sp.N, nt);
    size_t lds = (size_t)lds_slots(2 * sp.N) * sizeof(float2) + (size_t)(4 * sp.N) * sizeof(float);
    LP_DISPATCH_EPT(ept, k_sino_rows_pair, dim3((sp.P0 + 1) / 2, nz), dim3(nt), lds, s, R, Rc, sp, (int)cubic);
}

void launch_col_iir(const float* in, float* out, int nz, int L, int W, int chunk, int stuff, hipStream_t s) {
    dim3 block(256);
    switch (chunk) {
        case 32: hipLaunchKernelGGL(k_col_iir<32>, dim3((W + 255) / 256, (L + 31) / 32, nz), block, 0, s, in, out, L, W, stuff); break;
        case 128: hipLaunchKernelGGL(k_col_iir<128>, dim3((W + 255) / 256, (L + 127) / 128, nz), block, 0, s, in, out, L, W, stuff); break;
        default: hipLaunchKernelGGL(k_col_iir<64>, dim3((W + 255) / 256, (L + 63) / 64, nz), block, 0, s, in, out, L, W, stuff); break;
    }
}

void launch_gather(const Gather& g, const GTile* tiles, int ntiles, int kind, hipStream_t s) {
    if (ntiles <= 0) return;
    dim3 grid((ntiles / g.group) * ((g.nz + g.zgroup - 1) / g.zgroup));
    // whole 64-float chunks for stage_box
    size_t lds = (size_t)((g.box_floats + 63) & ~63) * sizeof(float);
    if (kind == 3 && g.group != 3) return;   // the Cartesian gather always sums three spans
    if (kind != 3 && g.group != 1) return;
    switch (kind) {
        case 0: hipLaunchKernelGGL((k_gather1<0>), grid, dim3(kGatherThreads), lds, s, g, tiles); break;
        case 1: hipLaunchKernelGGL((k_gather1<1>), grid, dim3(kGatherThreads), lds, s, g, tiles); break;
        case 2: hipLaunchKernelGGL((k_gather1<2>), grid, dim3(kGatherThreads), lds, s, g, tiles); break;
        default: hipLaunchKernelGGL((k_gather<3, 3>), grid, dim3(kGatherThreads), lds, s, g, tiles); break;
    }
}

// middle radix of h = 16 * MID * 16, or 0 when the register kernels do not apply
static int rows16_mid(const RowPass& rp) {
    int h = rp.h;
    if (!rp.fft16) return 0;
    if (h < 256 || h > 4096 || (h & (h - 1))) return 0;
    if ((long long)rp.R * (h / 16) > 1024) return 0;
    return h / 256;
}
// the 32-lane inverse row kernel: FFT kernel family 2, h = 1024, 16 rows per block
static bool rows32_applies(const RowPass& rp) { return rp.fft16 == 2 && rp.h == 1024 && rp.R == 16; }
static size_t rows16_lds(const RowPass& rp) {
    // the rows, then (h <= 2048) the staged twiddles twh[0, h/2) and twN[0, h/2]
    const size_t tws = rp.h <= 2048 ? (size_t)(rp.h + 1) * sizeof(float2) : 0;
    return (size_t)rp.R * (rp.h + rp.h / 16 + 32 / rp.R) * sizeof(float2) + tws + 64;
}

#define LP_DISPATCH_MID(mid, KER, grid, block, lds, stream, ...)                                                  \
    do {                                                                                                         \
        switch (mid) {                                                                                           \
            case 1: allow_lds(KER<1>, lds); hipLaunchKernelGGL(KER<1>, grid, block, lds, stream, __VA_ARGS__); break;   \
            case 2: allow_lds(KER<2>, lds); hipLaunchKernelGGL(KER<2>, grid, block, lds, stream, __VA_ARGS__); break;   \
            case 4: allow_lds(KER<4>, lds); hipLaunchKernelGGL(KER<4>, grid, block, lds, stream, __VA_ARGS__); break;   \
            case 8: allow_lds(KER<8>, lds); hipLaunchKernelGGL(KER<8>, grid, block, lds, stream, __VA_ARGS__); break;   \
            default: allow_lds(KER<16>, lds); hipLaunchKernelGGL(KER<16>, grid, block, lds, stream, __VA_ARGS__); break; \
        }                                                                                                        \
    } while (0)


// Slices per block, reduced for small batches so that a launch still has
// enough blocks to fill the chip several times over.
static RowPass fit_zgroup(const RowPass& in, long long blocks_per_slice) {
    RowPass rp = in;
    while (rp.zgroup > 1 && blocks_per_slice * ((rp.nz + rp.zgroup - 1) / rp.zgroup) < 2048) rp.zgroup /= 2;
    return rp;
}

void launch_rows_fwd(const RowPass& rp0, const float* F, const int2* range, hipStream_t s) {
    const RowPass rp = fit_zgroup(rp0, 3LL * (rp0.Nrho / rp0.R));
    if (int mid = rows16_mid(rp)) {
        dim3 grid((rp.Nrho / rp.R) * 3 * ((rp.nz + rp.zgroup - 1) / rp.zgroup));
        dim3 block(rp.R * (rp.h / 16));
        LP_DISPATCH_MID(mid, k_rows16_fwd, grid, block, rows16_lds(rp), s, rp, F, range);
        return;
    }
    // the mixed-radix kernels keep at most 8 rows per block (their register
    // budget); Nrho stays divisible since R is a power of two
    RowPass rg = rp;
    // the compiled pass has its rows per block built in
    rg.R = rg.rtc_rows_fwd ? rg.rtc_R : std::min(rg.R, 8);
    rg = fit_zgroup(rg, 3LL * (rg.Nrho / rg.R));
    dim3 grid(rg.Nrho / rg.R, 3, (rg.nz + rg.zgroup - 1) / rg.zgroup);
    if (rg.rtc_rows_fwd) {
        void* args[] = {&rg, &F, &range};
        if (hipModuleLaunchKernel((hipFunction_t)rg.rtc_rows_fwd, grid.x, grid.y, grid.z, rg.rtc_rows_nt, 1, 1, 0, s,
                                  args, nullptr) == hipSuccess)
            return;
        (void)hipGetLastError();   // the generic kernel does the pass
        RowPass rn = rp;
        rn.rtc_rows_fwd = nullptr;
        launch_rows_fwd(rn, F, range, s);
        return;
    }
    int nt = threads_for((long long)rg.R * rg.h);
    int ept = ept_for((long long)rg.R * rg.h, nt);
    int ldm = rg.ldm;
    LP_DISPATCH_EPT(ept, k_rows_fwd, grid, dim3(nt), rows_lds_bytes(ldm, rg.R), s, rg, F, range, ldm);
}

__global__ void k_add_extras(float* __restrict__ F, int nz, int Nt, int Nr, Src src, const int* __restrict__ xrow,
                             const float4* __restrict__ xent, int nx, int cubic) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int z = blockIdx.y;
    if (e >= nx) return;
    const int span = (e >= xrow[Nr + 1]) + (e >= xrow[2 * (Nr + 1)]);
    const float4 q = xent[e];
    const float v = sample(src.p + (size_t)z * src.zstride, src.W, src.H, q.y, q.z, cubic);
    atomicAdd(F + (((size_t)span * nz + z) * Nr + (int)q.w) * (size_t)Nt + (int)q.x, v);
}

void launch_add_extras(float* F, int nz, int Ntheta, int Nrho, const Src& src, const int* xrow, const float4* xent,
                       int nx, bool cubic, hipStream_t s) {
    if (!xrow || nx <= 0) return;
    hipLaunchKernelGGL(k_add_extras, dim3((nx + 255) / 256, nz), dim3(256), 0, s, F, nz, Ntheta, Nrho, src, xrow, xent,
                       nx, (int)cubic);
}

void launch_rows_inv(const RowPass& rp0, float* G, const int2* wrange, hipStream_t s) {
    const RowPass rp = fit_zgroup(rp0, 3LL * (rp0.Nrho / rp0.R));
    if (int mid = rows16_mid(rp)) {
        dim3 grid((rp.Nrho / rp.R) * 3 * ((rp.nz + rp.zgroup - 1) / rp.zgroup));
        if (rows32_applies(rp)) {
            allow_lds(k_rows32_inv, kRows32Lds);
            hipLaunchKernelGGL(k_rows32_inv, grid, dim3(512), kRows32Lds, s, rp, G, wrange);
            return;
        }
        dim3 block(rp.R * (rp.h / 16));
        LP_DISPATCH_MID(mid, k_rows16_inv, grid, block, rows16_lds(rp), s, rp, G, wrange);
        return;
    }
    RowPass rg = rp;
    rg.R = rg.rtc_rows_inv ? rg.rtc_R : std::min(rg.R, 8);
    rg = fit_zgroup(rg, 3LL * (rg.Nrho / rg.R));
    dim3 grid(rg.Nrho / rg.R, 3, (rg.nz + rg.zgroup - 1) / rg.zgroup);
    if (rg.rtc_rows_inv) {
        void* args[] = {&rg, &G, &wrange};
        if (hipModuleLaunchKernel((hipFunction_t)rg.rtc_rows_inv, grid.x, grid.y, grid.z, rg.rtc_rows_nt, 1, 1, 0, s,
                                  args, nullptr) == hipSuccess)
            return;
        (void)hipGetLastError();   // the generic kernel does the pass
        RowPass rn = rp;
        rn.rtc_rows_inv = nullptr;
        launch_rows_inv(rn, G, wrange, s);
        return;
    }
    int nt = threads_for((long long)rg.R * rg.h);
    int ept = ept_for((long long)rg.R * rg.h, nt);
    int ldm = rg.ldm;
    LP_DISPATCH_EPT(ept, k_rows_inv, grid, dim3(nt), rows_lds_bytes(ldm, rg.R), s, rg, G, wrange, ldm);
}

bool launch_rows_inv_polar(const RowPass& rp0, const PolarOut& po, hipStream_t s) {
    if (rp0.R != 16 || (rp0.Ntheta & (rp0.Ntheta - 1)) || (rp0.Nrho & (rp0.Nrho - 1))) return false;
    const RowPass rp = fit_zgroup(rp0, 3LL * po.nband);
    const int mid = rows16_mid(rp);
    if (!mid) return false;
    dim3 grid(po.nband * 3 * ((rp.nz + rp.zgroup - 1) / rp.zgroup));
    dim3 block(rp.R * (rp.h / 16));
    LP_DISPATCH_MID(mid, k_rows16_inv_polar, grid, block, rows16_lds(rp), s, rp, po);
    return true;
}

// middle radix of M = 16 * MID * 16, or 0 when the register kernel does not apply
static int cols16_mid(int M) {
    if (M < 256 || M > 4096 || (M & (M - 1))) return 0;
    return M / 256;
}

void launch_cols(const RowPass& rp0, const FftPlan& planM, const float2* twM, const float2* fzT, hipStream_t s) {
    const RowPass rp = fit_zgroup(rp0, 3LL * rp0.h);
    if (int mid = rp.fft16 ? cols16_mid(rp.Nrho) : 0) {
        const int T16 = rp.Nrho / 16;
        // column 0 (packed theta-DC / Nyquist pair) on the generic kernel, one
        // block per span and slice (a few blocks only: latency-bound)
        {
            RowPass r0 = rp;
            r0.zgroup = 1;
            int nt = threads_for(rp.Nrho);
            int ept = ept_for(rp.Nrho, nt);
            dim3 g0(1, 3, rp.nz);
            LP_DISPATCH_EPT(ept, k_cols, g0, dim3(nt), cols_lds_bytes(rp.Nrho), s, r0, planM, twM, fzT);
        }
        dim3 grid((rp.h - 1) * 3 * ((rp.nz + rp.zgroup - 1) / rp.zgroup));
        dim3 block(std::max(T16, 64));
        size_t lds = (size_t)lds_slots(rp.Nrho) * sizeof(float2);
        if (rp.h > 1) {
            switch (mid) {
                case 1: hipLaunchKernelGGL(k_cols16<1>, grid, block, lds, s, rp, twM, fzT); break;
                case 2: hipLaunchKernelGGL(k_cols16<2>, grid, block, lds, s, rp, twM, fzT); break;
                case 4: hipLaunchKernelGGL(k_cols16<4>, grid, block, lds, s, rp, twM, fzT); break;
                case 8: hipLaunchKernelGGL(k_cols16<8>, grid, block, lds, s, rp, twM, fzT); break;
                default: hipLaunchKernelGGL(k_cols4096, grid, block, 0, s, rp, twM, fzT); break;   // static LDS
            }
        }
        return;
    }
    dim3 grid(rp.h, 3, (rp.nz + rp.zgroup - 1) / rp.zgroup);
    if (rp.rtc_cols) {
        RowPass rc = rp;
        void* args[] = {&rc, &twM, &fzT};
        if (hipModuleLaunchKernel((hipFunction_t)rc.rtc_cols, grid.x, grid.y, grid.z, rc.rtc_cols_nt, 1, 1, 0, s, args,
                                  nullptr) == hipSuccess)
            return;
        (void)hipGetLastError();   // the generic kernel below does the pass
    }
    int nt = threads_for(rp.Nrho);
    int ept = ept_for(rp.Nrho, nt);
    LP_DISPATCH_EPT(ept, k_cols, grid, dim3(nt), cols_lds_bytes(rp.Nrho), s, rp, planM, twM, fzT);
}

// ---------------------------------------------------------------------------
// Transposes.  Each kernel below is the exact transpose of a stage above; the
// two chains fwd^T and adj^T run them in reverse stage order.

__global__ void k_zero(float4* __restrict__ p, size_t n4) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x)
        p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}
__global__ void k_zero_tail(float* __restrict__ p, size_t b, size_t n) {
    size_t i = b + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

// Transposed cubic B-spline prefilter of one line in LDS: the transpose of
// iir_line_lds.  The reference's recursion is y = A C x with C the causal
// sweep (12-term initialisation) and A the anticausal sweep; here
// u = A^T v is a causal sweep s[m] = v[m] + z s[m-1], u = -z s (z/(z-1) s at
// the end), and C^T u an anticausal sweep t[m] = u[m] + z t[m+1] giving
// lam t[m], plus lam z^(m+1) t[0] on m < 12 and lam (1+z) t[0] at m = 0.
__device__ void iir_line_lds_T(float* line, float* tmp, int L) {
    const float z = kPole;
    const float kap = z / (z - 1.0f);
    const int nt = blockDim.x;
    const int T = ((L + nt - 1) / nt) | 1;
    const int s = threadIdx.x * T;
    const int e = min(s + T, L);
    if (s < L) {
        int a = max(0, s - kHalo);   // zero before row 0, so a = 0 is exact
        float c = line[a];
        if (a >= s) tmp[a] = (a == L - 1) ? kap * c : -z * c;
        for (int n = a + 1; n < e; ++n) {
            c = line[n] + z * c;
            if (n >= s) tmp[n] = (n == L - 1) ? kap * c : -z * c;
        }
    }
    __syncthreads();
    if (s < L) {
        int b = min(e - 1 + kHalo, L - 1);
        float c = tmp[b];
        if (b < e) line[b] = kLambda * c;
        for (int n = b - 1; n >= s; --n) {
            c = tmp[n] + z * c;
            if (n < e) line[n] = kLambda * c;
        }
    }
    __syncthreads();
    const float l0 = line[0];   // lam t[0]
    __syncthreads();
    const int hz = min(kHorizon, L);
    for (int m = threadIdx.x; m < hz; m += nt) {
        float zp = z;
        for (int q = 0; q < m; ++q) zp *= z;   // z^(m+1)
        line[m] = (m == 0) ? (1.0f + z) * l0 : line[m] + zp * l0;
    }
    __syncthreads();
}

// Transpose of k_col_iir: same T-row chunks with a kHalo halo on each side.
template <int T>
__global__ void k_col_iir_T(const float* __restrict__ in, float* __restrict__ outp, int L, int W) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int r0 = blockIdx.y * T;
    const int z = blockIdx.z;
    if (x >= W) return;
    constexpr int K = kHalo;
    constexpr int NV = T + 2 * K;
    const float* col = in + (size_t)z * L * W + x;
    float* ocol = outp + (size_t)z * L * W + x;
    const float zp = kPole;
    const float kap = zp / (zp - 1.0f);
    float v[NV];
#pragma unroll
    for (int n = 0; n < NV; ++n) {
        int row = r0 - K + n;
        v[n] = (row >= 0 && row < L) ? col[(size_t)row * W] : 0.f;
    }
    // transposed anticausal stage: causal sweep (zeros before row 0 make the
    // first chunk exact), u = -z s, kap s at row L-1, nothing past it
    float c = 0.f;
#pragma unroll
    for (int n = 0; n < NV; ++n) {
        int row = r0 - K + n;
        c = v[n] + zp * c;
        v[n] = (row < L - 1) ? -zp * c : (row == L - 1 ? kap * c : 0.f);
    }
    // transposed causal stage: anticausal sweep from the top of the window
    c = 0.f;
#pragma unroll
    for (int n = NV - 1; n >= 0; --n) {
        c = v[n] + zp * c;
        v[n] = c;
    }
    if (r0 == 0) {
        // the causal initialisation's column: + z^(m+1) t[0] on m < 12, so
        // that row 0 holds (1 + z) t[0]
        const float t0 = v[K];
        const int hz = min(kHorizon, L);
        float zq = zp;
#pragma unroll
        for (int m = 0; m < kHorizon; ++m) {
            if (m < hz) v[K + m] += zq * t0;
            zq *= zp;
        }
    }
#pragma unroll
    for (int n = K; n < K + T; ++n) {
        int row = r0 - K + n;
        if (row < L) ocol[(size_t)row * W] = kLambda * v[n];
    }
}

// Transpose of k_image_rows: x-direction prefilter transpose, then the crop
// back to the user's N0 x N0 image (the transpose of the zero padding).
__global__ void k_image_rows_T(const float* __restrict__ fc, float* __restrict__ f, int N0, int N, int o, int cubic) {
    extern __shared__ float sm[];
    float* line = sm;
    float* tmp = sm + N;
    const int yi = blockIdx.x + o, z = blockIdx.y;
    const float* src = fc + ((size_t)z * N + yi) * N;
    for (int x = threadIdx.x; x < N; x += blockDim.x) line[x] = src[x];
    __syncthreads();
    if (cubic) iir_line_lds_T(line, tmp, N);
    float* dst = f + ((size_t)z * N0 + (yi - o)) * N0;
    for (int x = threadIdx.x; x < N0; x += blockDim.x) dst[x] = line[x + o];
}

// Transpose of k_sino_rows for the kept angle rows: s-direction prefilter
// transpose, the FBP convolution (a symmetric Toeplitz matrix: its own
// transpose), then the transpose of the edge replication (the columns left of
// lb fold into column lb, those right of rb into rb) and of the placement.
template <int EPT, bool FILT>
__global__ void k_sino_rows_T(const float* __restrict__ Rc, float* __restrict__ R, SinoPrep sp, int cubic) {
    extern __shared__ float2 smc[];
    __shared__ float red[2][32];
    const int N = sp.N;
    const int a = blockIdx.x, z = blockIdx.y;
    const int ai = a * sp.os;
    const float* src = Rc + ((size_t)z * sp.P + ai) * N;
    float* line = FILT ? reinterpret_cast<float*>(smc + lds_slots(sp.Lh)) : reinterpret_cast<float*>(smc);
    float* tmp = line + N;
    for (int x = threadIdx.x; x < N; x += blockDim.x) line[x] = src[x];
    __syncthreads();
    if (cubic) iir_line_lds_T(line, tmp, N);
    if (FILT) {
        // the filter is symmetric (a real even spectrum): its own transpose
        filter_line<EPT>(smc, line, sp);
        __syncthreads();
    }
    // sums of the replicated columns [0, lb) and (rb, N)
    const int lb = sp.lb, rb = sp.lb + sp.N0 - 1;
    float sl = 0.f, sr = 0.f;
    for (int x = threadIdx.x; x < N; x += blockDim.x) {
        if (x < lb) sl += line[x];
        else if (x > rb) sr += line[x];
    }
    for (int off = 32; off > 0; off >>= 1) {
        sl += __shfl_down(sl, off);
        sr += __shfl_down(sr, off);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) { red[0][wid] = sl; red[1][wid] = sr; }
    __syncthreads();
    if (threadIdx.x == 0) {
        float tl = 0.f, tr = 0.f;
        for (int w = 0; w < (int)(blockDim.x + 63) / 64; ++w) { tl += red[0][w]; tr += red[1][w]; }
        line[lb] += tl;
        line[rb] += tr;
    }
    __syncthreads();
    float* dst = R + ((size_t)z * sp.P0 + a) * sp.N0;
    for (int x = threadIdx.x; x < sp.N0; x += blockDim.x) dst[x] = line[x + lb];
}

// Transpose of k_gather: per tile, the block clears an LDS copy of the source
// box, adds every point's value times its tap weights into it (LDS atomics),
// and adds the box into the destination grid (global atomics: neighbouring
// boxes overlap).  Same tiles, logical block order and clamps as k_gather.
template <int KIND>
__global__ void __launch_bounds__(kGatherThreads) k_scatter(Scatter g, const GTile* __restrict__ tiles) {
    extern __shared__ float box[];
    constexpr int PPT = kGatherPPT;
    const int nt = blockDim.x;
    const int nzg = (g.nz + g.zgroup - 1) / g.zgroup;
    const int lbk = xcd_block(blockIdx.x, gridDim.x);
    const int ng = gridDim.x / nzg;
    const int zg = lbk % nzg, tile = (int)(((long long)(lbk / nzg) * g.tstride) % ng);
    const GTile* tb = tiles + (size_t)tile * g.group;
    const GTile t0 = tb[0];
    const int npts = t0.nr * t0.nc;
    const int z0 = zg * g.zgroup, z1 = min(z0 + g.zgroup, g.nz);
    if (g.perm) {   // k_scatter_col handles the coloured tiles
        bool any = false;
        for (int gi = 0; gi < g.group; ++gi) any |= tb[gi].bw > 0 && tb[gi].cr == 0;
        if (!any) return;
    }
    // Loads are unconditional at clamped point indices (a guarded load
    // compiles to a branch and a wait of its own), all PPT in flight at once.
    const float inv_nc0 = 1.0f / (float)t0.nc;
    for (int z = z0; z < z1; ++z) {
        float val[PPT];
        const float* V = g.val + (size_t)z * g.v_z + (size_t)t0.span * g.v_span;
#pragma unroll
        for (int e = 0; e < PPT; ++e) {
            int r, c;
            tile_rc(min(threadIdx.x + e * nt, npts - 1), t0.nc, inv_nc0, r, c);
            val[e] = V[(size_t)(t0.r0 + r) * g.v_row + t0.c0 + c];
        }
        if (g.scale) {   // uniform
#pragma unroll
            for (int e = 0; e < PPT; ++e) {
                int r, c;
                tile_rc(min(threadIdx.x + e * nt, npts - 1), t0.nc, inv_nc0, r, c);
                val[e] *= g.scale[(size_t)(t0.r0 + r) * g.s_row + t0.c0 + c];
            }
        }
#pragma unroll
        for (int e = 0; e < PPT; ++e)
            if (threadIdx.x + e * nt >= npts) val[e] = 0.f;
        for (int gi = 0; gi < g.group; ++gi) {
            const GTile t = tb[gi];
            if (t.bw <= 0 || (g.perm && t.cr > 0)) continue;   // uniform over the block
            const int bs = t.bs;
            const int nbox = bs * t.bh;
            const float2* cbase = g.coord + (size_t)t.span * g.c_span;
            float2 cc[PPT];
#pragma unroll
            for (int e = 0; e < PPT; ++e) {
                int r, cidx;
                tile_rc(min(threadIdx.x + e * nt, npts - 1), t.nc, 1.0f / (float)t.nc, r, cidx);
                cc[e] = cbase[(size_t)(t.r0 + r + g.c_ro) * g.c_row + t.c0 + cidx + g.c_co];
            }
            for (int i = threadIdx.x; i < nbox; i += nt) box[i] = 0.f;
            __syncthreads();
            // one point's 16 (4) weighted taps into the box, atomically
            auto deposit = [&](int e, bool atomic) {
                const float2 c = cc[e];
                float fx = floorf(c.x), fy = floorf(c.y);
                float ax = c.x - fx, ay = c.y - fy;
                int lx = min(max((int)fx - 1 - t.x0, 0), t.bw - 4);
                int ly = min(max((int)fy - 1 - t.y0, 0), t.bh - 4);
                float* q = box + ly * bs + lx;
                const float v = val[e];
                auto add = [&](float* a, float w) {
                    if (atomic) atomicAdd(a, w);
                    else *a += w;
                };
                if (g.cubic) {
                    float wx[4], wy[4];
                    bspline_w(ax, wx);
                    bspline_w(ay, wy);
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        float vb = wy[b] * v;
#pragma unroll
                        for (int a2 = 0; a2 < 4; ++a2) add(q + b * bs + a2, wx[a2] * vb);
                    }
                } else {
                    float* q1 = q + bs + 1;
                    float* q2 = q1 + bs;
                    add(q1, (1.f - ay) * (1.f - ax) * v);
                    add(q1 + 1, (1.f - ay) * ax * v);
                    add(q2, ay * (1.f - ax) * v);
                    add(q2 + 1, ay * ax * v);
                }
            };
            {
#pragma unroll
                for (int e = 0; e < PPT; ++e) {
                    int p = threadIdx.x + e * nt;
                    if (p < npts && val[e] != 0.f && cc[e].x == cc[e].x) deposit(e, true);
                }
                __syncthreads();
            }
            const float inv = 1.0f / (float)t.bw;
            const int nb = t.bw * t.bh;
            if (g.stage) {
                float* B = g.stage + (size_t)z * g.stage_z + g.box_off[(size_t)tile * g.group + gi];
                for (int i = threadIdx.x; i < nb; i += nt) {
                    int yy = (int)(((float)i + 0.5f) * inv);
                    int xx = i - yy * t.bw;
                    B[i] = box[yy * bs + xx];
                }
            } else {
                float* D = g.dst + (size_t)z * g.dst_z + (size_t)t.span * g.dst_span;
                for (int i = threadIdx.x; i < nb; i += nt) {
                    int yy = (int)(((float)i + 0.5f) * inv);
                    int xx = i - yy * t.bw;
                    float v = box[yy * bs + xx];
                    if (v != 0.f) unsafeAtomicAdd(D + (size_t)wrapi(t.y0 + yy, g.H) * g.W + wrapi(t.x0 + xx, g.W), v);
                }
            }
            __syncthreads();
        }
    }
}

// Transpose of the extra samples added in the row pass: block (row, span, z)
// scatters F[span][z][row][theta] of each of its extra entries into dst.
__global__ void k_scatter_extra(const float* __restrict__ F, float* __restrict__ dst, int nz, int Nt, int Nr, int W,
                                int H, const int* __restrict__ xrow, const float4* __restrict__ xent, int cubic) {
    const int row = blockIdx.x, span = blockIdx.y, z = blockIdx.z;
    const int lo = xrow[span * (Nr + 1) + row], hi = xrow[span * (Nr + 1) + row + 1];
    const float* Fr = F + (((size_t)span * nz + z) * Nr + row) * (size_t)Nt;
    float* D = dst + (size_t)z * W * H;
    for (int e = lo + threadIdx.x; e < hi; e += blockDim.x) {
        float4 q = xent[e];
        float v = Fr[(int)q.x];
        float x = q.y, y = q.z;
        float fx = floorf(x), fy = floorf(y);
        float ax = x - fx, ay = y - fy;
        int ix = (int)fx, iy = (int)fy;
        if (cubic) {
            float wx[4], wy[4];
            bspline_w(ax, wx);
            bspline_w(ay, wy);
            for (int b = 0; b < 4; ++b)
                for (int a = 0; a < 4; ++a)
                    unsafeAtomicAdd(D + (size_t)wrapi(iy - 1 + b, H) * W + wrapi(ix - 1 + a, W), wy[b] * wx[a] * v);
        } else {
            for (int b = 0; b < 2; ++b)
                for (int a = 0; a < 2; ++a)
                    unsafeAtomicAdd(D + (size_t)wrapi(iy + b, H) * W + wrapi(ix + a, W),
                                    (b ? ay : 1.f - ay) * (a ? ax : 1.f - ax) * v);
        }
    }
}

void launch_zero(float* p, size_t n, hipStream_t s) {
    size_t n4 = n / 4;
    if (n4) {
        size_t blocks = std::min<size_t>((n4 + 255) / 256, 65536);
        hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, s, reinterpret_cast<float4*>(p), n4);
    }
    if (n % 4) hipLaunchKernelGGL(k_zero_tail, dim3(1), dim3(4), 0, s, p, n4 * 4, n);
}

// The tiled scatter for coloured tiles: the tile's points are taken colour
// class by colour class (host-sorted), one point per lane; the points of one
// class have disjoint tap footprints, so each box word takes at most one
// plain read-modify-write per class -- no atomics, and the order of the
// additions is fixed.  Same block order, boxes and flush as k_scatter.
template <int KIND>
__global__ void __launch_bounds__(kGatherThreads) k_scatter_col(Scatter g, const GTile* __restrict__ tiles) {
    extern __shared__ float box[];
    // KZ slices share each pass over the colour classes: a point's tap
    // geometry is computed once and added into KZ boxes (one per slice, LDS
    // KZ * box_floats), so the block has half the class barriers per slice
    constexpr int KZ = kScatterSlices;
    const int nt = blockDim.x;
    const int nzg = (g.nz + g.zgroup - 1) / g.zgroup;
    const int lbk = xcd_block(blockIdx.x, gridDim.x);
    const int ng = gridDim.x / nzg;
    const int zg = lbk % nzg, tile = (int)(((long long)(lbk / nzg) * g.tstride) % ng);
    const GTile* tb = tiles + (size_t)tile * g.group;
    const GTile t0 = tb[0];
    const int z0 = zg * g.zgroup, z1 = min(z0 + g.zgroup, g.nz);
    const int bstride = g.box_floats;
    for (int zb = z0; zb < z1; zb += KZ) {
        const int zn = min(KZ, z1 - zb);
        const float* V = g.val + (size_t)zb * g.v_z + (size_t)t0.span * g.v_span;
        for (int gi = 0; gi < g.group; ++gi) {
            const GTile t = tb[gi];
            if (t.bw <= 0 || t.cr == 0) continue;   // uniform over the block
            const int bs = t.bs;
            const int nbox = bs * t.bh;
            const float2* cb = g.coord + (size_t)t.span * g.c_span;
            for (int q = 0; q < zn; ++q)
                for (int i = threadIdx.x; i < nbox; i += nt) box[q * bstride + i] = 0.f;
            __syncthreads();
            const int ncol = t.cr * t.cc;
            for (int col = 0; col < ncol; ++col) {
                const int pb = g.cls[t.cbase + col], pe = g.cls[t.cbase + col + 1];
#pragma unroll 1
                for (int i = pb + (int)threadIdx.x; i < pe; i += nt) {
                    const int p = g.perm[t.pbase + i];
                    const int r = p >> t.lnc, cidx = p & (t.nc - 1);
                    const float2 c = cb[(size_t)(t.r0 + r + g.c_ro) * g.c_row + t.c0 + cidx + g.c_co];
                    const size_t vo = (size_t)(t0.r0 + r) * g.v_row + t0.c0 + cidx;
                    const float sc = g.scale ? g.scale[(size_t)(t0.r0 + r) * g.s_row + t0.c0 + cidx] : 1.f;
                    float v[KZ];
#pragma unroll
                    for (int q = 0; q < KZ; ++q) v[q] = q < zn ? V[vo + (size_t)q * g.v_z] : 0.f;
                    if (g.scale) {
#pragma unroll
                        for (int q = 0; q < KZ; ++q) v[q] *= sc;
                    }
                    const float fx = floorf(c.x), fy = floorf(c.y);
                    const float ax = c.x - fx, ay = c.y - fy;
                    const int lx = min(max((int)fx - 1 - t.x0, 0), t.bw - 4);
                    const int ly = min(max((int)fy - 1 - t.y0, 0), t.bh - 4);
                    if (g.cubic) {
                        float wx[4], wy[4];
                        bspline_w(ax, wx);
                        bspline_w(ay, wy);
#pragma unroll
                        for (int q = 0; q < KZ; ++q) {
                            if (q >= zn) break;
                            float* qb = box + q * bstride + ly * bs + lx;
#pragma unroll
                            for (int b = 0; b < 4; ++b) {
                                const float vb = wy[b] * v[q];
                                float* qr = qb + b * bs;
                                const float s0 = qr[0], s1 = qr[1], s2 = qr[2], s3 = qr[3];
                                qr[0] = s0 + wx[0] * vb;
                                qr[1] = s1 + wx[1] * vb;
                                qr[2] = s2 + wx[2] * vb;
                                qr[3] = s3 + wx[3] * vb;
                            }
                        }
                    } else {
#pragma unroll
                        for (int q = 0; q < KZ; ++q) {
                            if (q >= zn) break;
                            float* q1 = box + q * bstride + ly * bs + lx + bs + 1;
                            float* q2 = q1 + bs;
                            q1[0] += (1.f - ay) * (1.f - ax) * v[q];
                            q1[1] += (1.f - ay) * ax * v[q];
                            q2[0] += ay * (1.f - ax) * v[q];
                            q2[1] += ay * ax * v[q];
                        }
                    }
                }
                __syncthreads();
            }
            const float inv = 1.0f / (float)t.bw;
            const int nb = t.bw * t.bh;
            for (int q = 0; q < zn; ++q) {
                const float* bq = box + q * bstride;
                if (g.stage) {
                    float* B = g.stage + (size_t)(zb + q) * g.stage_z + g.box_off[(size_t)tile * g.group + gi];
                    for (int i = threadIdx.x; i < nb; i += nt) {
                        const int yy = (int)(((float)i + 0.5f) * inv);
                        B[i] = bq[yy * bs + (i - yy * t.bw)];
                    }
                } else {
                    float* D = g.dst + (size_t)(zb + q) * g.dst_z + (size_t)t.span * g.dst_span;
                    for (int i = threadIdx.x; i < nb; i += nt) {
                        const int yy = (int)(((float)i + 0.5f) * inv);
                        const int xx = i - yy * t.bw;
                        const float v = bq[yy * bs + xx];
                        if (v != 0.f) unsafeAtomicAdd(D + (size_t)wrapi(t.y0 + yy, g.H) * g.W + wrapi(t.x0 + xx, g.W), v);
                    }
                }
            }
            __syncthreads();
        }
    }
}

void launch_scatter(const Scatter& g0, const GTile* tiles, int ntiles, int kind, hipStream_t s) {
    if (ntiles <= 0) return;
    Scatter g = g0;
    // tile order: a stride coprime with the number of tile groups, about 0.62
    // of it, so that blocks running together come from far apart in the tile
    // list (neighbouring tiles flush into the same destination pixels)
    const int ng = ntiles / g.group;
    int st = std::max(1, (int)(0.618 * ng));
    while (std::gcd(st, ng) != 1) ++st;
    g.tstride = st;
    dim3 grid(ng * ((g.nz + g.zgroup - 1) / g.zgroup));
    size_t lds = (size_t)g.box_floats * sizeof(float);
    if (g.perm) {
        const size_t ldsc = lds * kScatterSlices;
        switch (kind) {
            case 0: hipLaunchKernelGGL(k_scatter_col<0>, grid, dim3(kGatherThreads), ldsc, s, g, tiles); break;
            case 1: hipLaunchKernelGGL(k_scatter_col<1>, grid, dim3(kGatherThreads), ldsc, s, g, tiles); break;
            case 2: hipLaunchKernelGGL(k_scatter_col<2>, grid, dim3(kGatherThreads), ldsc, s, g, tiles); break;
            default: hipLaunchKernelGGL(k_scatter_col<3>, grid, dim3(kGatherThreads), ldsc, s, g, tiles); break;
        }
    }
    switch (kind) {
        case 0: hipLaunchKernelGGL(k_scatter<0>, grid, dim3(kGatherThreads), lds, s, g, tiles); break;
        case 1: hipLaunchKernelGGL(k_scatter<1>, grid, dim3(kGatherThreads), lds, s, g, tiles); break;
        case 2: hipLaunchKernelGGL(k_scatter<2>, grid, dim3(kGatherThreads), lds, s, g, tiles); break;
        default: hipLaunchKernelGGL(k_scatter<3>, grid, dim3(kGatherThreads), lds, s, g, tiles); break;
    }
}

// Step A of the polar transpose: a block of 256 rows of one (span, angle);
// the rows' detector runs are consecutive, so the coordinate and value loads
// of neighbouring lanes are too.  u [nz][3][nA][Nrho].
__global__ void __launch_bounds__(256) k_polarT_rows(PolarT p) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int al = blockIdx.y, k = blockIdx.z;
    if (r >= p.Nr) return;
    const int a0 = p.abeg[k], na = p.abeg[3 + k] - a0;
    float* U = p.u + ((size_t)k * p.nA + al) * p.Nr + r;
    const size_t u_z = (size_t)3 * p.nA * p.Nr;
    if (al >= na) {
        for (int z = 0; z < p.nz; ++z) U[z * u_z] = 0.f;
        return;
    }
    const int a = a0 + al;
    const int* xa = p.xs + (size_t)a * (p.Q + 1);
    const float2* ca = p.coord + (size_t)a * p.c_row + p.c_co;
    // unwrapped rows q = r (mod Nrho) inside the table's range; an angle's
    // taps wrap around rho at most once, so at most two q have cells
    int x0[2], x1[2], qq[2], nq = 0;
    for (int q = p.qlo + (((r - p.qlo) % p.Nr) + p.Nr) % p.Nr; q <= p.qlo + p.Q && nq < 2; q += p.Nr) {
        const int i = q - p.qlo;
        const int lo = p.cubic ? i - 2 : i - 1, hi = p.cubic ? i + 2 : i + 1;
        const int b0 = xa[max(lo, 0)], b1 = xa[min(hi, p.Q)];
        if (b0 < b1) { x0[nq] = b0; x1[nq] = b1; qq[nq] = q; ++nq; }
    }
    for (int z = 0; z < p.nz; ++z) {
        const float* V = p.val + (size_t)z * p.v_z + (size_t)a * p.N0;
        float acc = 0.f;
        for (int t = 0; t < nq; ++t)
            for (int x = x0[t]; x < x1[t]; ++x) {
                const float y = ca[x].y;
                const float fy = floorf(y);
                const int b = qq[t] - (int)fy + (p.cubic ? 1 : 0);
                const float ay = y - fy;
                float w;
                if (p.cubic) {
                    float wy[4];
                    bspline_w(ay, wy);
                    w = b == 0 ? wy[0] : b == 1 ? wy[1] : b == 2 ? wy[2] : wy[3];
                } else {
                    w = b == 0 ? 1.f - ay : ay;
                }
                acc += w * V[x];
            }
        U[z * u_z] = acc;
    }
}

// Step B: a block owns 64 columns x 64 rows of one span; the u rows of the
// angles whose taps reach those columns are staged through LDS 128 angles at
// a time, and each lane sums its column for 16 rows.
__global__ void __launch_bounds__(256) k_polarT_cols(PolarT p) {
    __shared__ float us[128][65];
    const int k = blockIdx.z, r0 = blockIdx.y * 64, jj0 = blockIdx.x * 64;
    const int lc = threadIdx.x & 63, lr = threadIdx.x >> 6;
    const int jj = jj0 + lc, j = p.jlo + jj;
    const int* A = p.aj + (size_t)k * (p.jn + 5);
    const int jl = min(jj, p.jn - 1);
    const int lo = jj < p.jn ? (p.cubic ? A[jl] : A[jl + 1]) : 0;
    const int hi = jj < p.jn ? (p.cubic ? A[jl + 4] : A[jl + 3]) : 0;
    const int jlast = min(jj0 + 63, p.jn - 1);
    const int blo = p.cubic ? A[jj0] : A[jj0 + 1], bhi = p.cubic ? A[jlast + 4] : A[jlast + 3];
    const int a0 = p.abeg[k];
    const size_t u_z = (size_t)3 * p.nA * p.Nr;
    const size_t g_z = (size_t)p.Nr * p.Nt;
    for (int z = 0; z < p.nz; ++z) {
        float acc[16];
#pragma unroll
        for (int m = 0; m < 16; ++m) acc[m] = 0.f;
        const float* U = p.u + z * u_z + ((size_t)k * p.nA) * p.Nr;
        for (int c0 = blo; c0 < bhi; c0 += 128) {
            const int cn = min(128, bhi - c0);
            __syncthreads();
            for (int i = threadIdx.x; i < cn * 64; i += blockDim.x) {
                const int ai = i >> 6, ri = i & 63;
                const int rr = r0 + ri;
                us[ai][ri] = rr < p.Nr ? U[(size_t)(c0 + ai) * p.Nr + rr] : 0.f;
            }
            __syncthreads();
            for (int al = max(lo, c0); al < min(hi, c0 + cn); ++al) {
                const float x = p.cxa[a0 + al];
                const float fx = floorf(x);
                const int c = j - (int)fx + (p.cubic ? 1 : 0);
                const float ax = x - fx;
                float w;
                if (p.cubic) {
                    float wx[4];
                    bspline_w(ax, wx);
                    w = c == 0 ? wx[0] : c == 1 ? wx[1] : c == 2 ? wx[2] : wx[3];
                } else {
                    w = c == 0 ? 1.f - ax : ax;
                }
#pragma unroll
                for (int m = 0; m < 16; ++m) acc[m] += w * us[al - c0][lr + 4 * m];
            }
        }
        if (jj < p.jn) {
            float* Gk = p.G + ((size_t)k * p.nz + z) * g_z + j;
#pragma unroll
            for (int m = 0; m < 16; ++m) {
                const int rr = r0 + lr + 4 * m;
                if (rr < p.Nr) Gk[(size_t)rr * p.Nt] = acc[m];
            }
        }
    }
}

void launch_polar_transpose(const PolarT& p, hipStream_t s) {
    hipLaunchKernelGGL(k_polarT_rows, dim3((p.Nr + 255) / 256, p.nA, 3), dim3(256), 0, s, p);
    hipLaunchKernelGGL(k_polarT_cols, dim3((p.jn + 63) / 64, (p.Nr + 63) / 64, 3), dim3(256), 0, s, p);
}

// One block per destination tile of 32 x 32 points (4 per thread); the
// tile's box list is walked once per 8 slices, its rectangles and offsets
// staged in LDS 128 at a time (one burst of list reads instead of a
// dependent pair of global loads per box), and two boxes' value loads are in
// flight together.  Every point adds its boxes in list order.
__global__ void __launch_bounds__(256) k_box_sum(BoxSum b) {
    constexpr int ZC = 8, LB = 128;
    __shared__ int4 srect[LB];
    __shared__ int soff[LB];
    const int t = blockIdx.x;
    const int plane = t / (b.tiles_x * b.tiles_y);
    const int tt = t - plane * b.tiles_x * b.tiles_y;
    const int ty = tt / b.tiles_x, tx = tt - ty * b.tiles_x;
    const int X = tx * 32 + (threadIdx.x & 31);
    const int Y0 = ty * 32 + (threadIdx.x >> 5);   // rows Y0 + 8 e, e < 4
    const int beg = b.tile_beg[t], end = b.tile_beg[t + 1];
    for (int z0 = 0; z0 < b.nz; z0 += ZC) {
        const int zn = min(ZC, b.nz - z0);
        float acc[ZC][4];
#pragma unroll
        for (int q = 0; q < ZC; ++q)
#pragma unroll
            for (int e = 0; e < 4; ++e) acc[q][e] = 0.f;
        const float* St = b.stage + (size_t)z0 * b.stage_z;
        for (int c0 = beg; c0 < end; c0 += LB) {
            const int nb = min(LB, end - c0);
            __syncthreads();   // the previous round's entries are read
            if ((int)threadIdx.x < nb) {
                const int id = b.tile_box[c0 + threadIdx.x];
                srect[threadIdx.x] = b.box_rect[id];
                soff[threadIdx.x] = b.box_off[id];
            }
            __syncthreads();
#pragma unroll 2
            for (int i = 0; i < nb; ++i) {
                const int4 r = srect[i];
                const float* B = St + soff[i];
                // box origins lie within one period of the plane: one wrap at most
                int dx = X - r.x;
                dx += dx < 0 ? b.W : (dx >= b.W ? -b.W : 0);
                if (dx >= r.z) continue;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    int dy = Y0 + 8 * e - r.y;
                    dy += dy < 0 ? b.H : (dy >= b.H ? -b.H : 0);
                    if (dy < r.w) {
                        const float* Bq = B + dy * r.z + dx;
#pragma unroll
                        for (int q = 0; q < ZC; ++q)
                            if (q < zn) acc[q][e] += Bq[(size_t)q * b.stage_z];
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < ZC; ++q) {
            if (q >= zn) break;
            float* D = b.dst + (size_t)(z0 + q) * b.dst_z + (size_t)plane * b.dst_plane;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int Y = Y0 + 8 * e;
                if (X < b.W && Y < b.H) D[(size_t)Y * b.W + X] = acc[q][e];
            }
        }
    }
}

void launch_box_sum(const BoxSum& b, hipStream_t s) {
    const int nt = b.nplanes * b.tiles_x * b.tiles_y;
    if (nt > 0) hipLaunchKernelGGL(k_box_sum, dim3(nt), dim3(256), 0, s, b);
}

void launch_scatter_extra(const float* F, float* dst, int nz, int Ntheta, int Nrho, int W, int H, const int* xrow,
                          const float4* xent, bool cubic, hipStream_t s) {
    if (!xrow) return;
    hipLaunchKernelGGL(k_scatter_extra, dim3(Nrho, 3, nz), dim3(64), 0, s, F, dst, nz, Ntheta, Nrho, W, H, xrow, xent,
                       (int)cubic);
}

void launch_col_iir_T(const float* in, float* out, int nz, int L, int W, int chunk, hipStream_t s) {
    dim3 block(256);
    switch (chunk) {
        case 32: hipLaunchKernelGGL(k_col_iir_T<32>, dim3((W + 255) / 256, (L + 31) / 32, nz), block, 0, s, in, out, L, W); break;
        case 128: hipLaunchKernelGGL(k_col_iir_T<128>, dim3((W + 255) / 256, (L + 127) / 128, nz), block, 0, s, in, out, L, W); break;
        default: hipLaunchKernelGGL(k_col_iir_T<64>, dim3((W + 255) / 256, (L + 63) / 64, nz), block, 0, s, in, out, L, W); break;
    }
}

void launch_image_rows_T(const float* fc, float* f, int nz, int N0, int N, bool cubic, hipStream_t s) {
    hipLaunchKernelGGL(k_image_rows_T, dim3(N0, nz), dim3(256), 2 * N * sizeof(float), s, fc, f, N0, N,
                       N / 2 - N0 / 2, (int)cubic);
}

void launch_sino_rows_T(const float* Rc, float* R, int nz, const SinoPrep& sp, bool cubic, hipStream_t s) {
    if (!sp.H2) {
        hipLaunchKernelGGL((k_sino_rows_T<4, false>), dim3(sp.P0, nz), dim3(256), 2 * (size_t)sp.N * sizeof(float), s,
                           Rc, R, sp, (int)cubic);
        return;
    }
    int nt = threads_for(sp.Lh);
    int ept = ept_for(sp.Lh, nt);
    LP_DISPATCH_EPT2(ept, k_sino_rows_T, true, dim3(sp.P0, nz), dim3(nt), sino_lds_bytes(sp.Lh, sp.N), s, Rc, R, sp,
                     (int)cubic);
}

// ---------------------------------------------------------------------------
// Streaming kernels that measure the HBM ceiling the pipeline's passes are
// compared with: 16-byte loads and stores, UL independent loads in flight per
// lane, grid-stride over n4 float4.  kind 0 copies (read + write), 1 reads
// (the per-lane sums land in dst[thread], so the loads cannot be dropped),
// 2 writes.
typedef float v4f __attribute__((ext_vector_type(4)));

template <int KIND>
__global__ void __launch_bounds__(256) k_stream(v4f* __restrict__ dst, const v4f* __restrict__ src, size_t n4) {
    constexpr int UL = 4;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    const v4f one = {1.f, 1.f, 1.f, 1.f};
    for (; i + (UL - 1) * stride < n4; i += UL * stride) {
        v4f v[UL];
        if (KIND != 2) {
#pragma unroll
            for (int u = 0; u < UL; ++u) v[u] = __builtin_nontemporal_load(src + i + u * stride);
        }
#pragma unroll
        for (int u = 0; u < UL; ++u) {
            if (KIND == 0) __builtin_nontemporal_store(v[u], dst + i + u * stride);
            else if (KIND == 1) acc += v[u];
            else __builtin_nontemporal_store(one, dst + i + u * stride);
        }
    }
    for (; i < n4; i += stride) {
        if (KIND == 0) dst[i] = src[i];
        else if (KIND == 1) acc += src[i];
        else dst[i] = one;
    }
    if (KIND == 1) dst[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

int launch_stream(int kind, float* dst, const float* src, size_t n4, int blocks, hipStream_t s) {
    if (blocks < 1) blocks = 2048;
    v4f* d = reinterpret_cast<v4f*>(dst);
    const v4f* a = reinterpret_cast<const v4f*>(src);
    switch (kind) {
        case 0: hipLaunchKernelGGL(k_stream<0>, dim3(blocks), dim3(256), 0, s, d, a, n4); return 0;
        case 1: hipLaunchKernelGGL(k_stream<1>, dim3(blocks), dim3(256), 0, s, d, a, n4); return 0;
        case 2: hipLaunchKernelGGL(k_stream<2>, dim3(blocks), dim3(256), 0, s, d, a, n4); return 0;
    }
    return -1;
}

}  // namespace lpk
