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
#include "kernels.hpp"

namespace lpk {

__device__ __forceinline__ int wrapi(int i, int n) {
    if ((unsigned)i < (unsigned)n) return i;
    i %= n;
    return i < 0 ? i + n : i;
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

// Cubic B-spline prefilter of one line held in LDS (ConvertToInterpolationCoefficients,
// main_kernels.cuh:97-142).  Each thread owns a chunk and restarts the causal and
// anticausal recursions kHalo samples outside it (truncation |z|^17 < 2e-10), except
// at the line ends where the reference's own initialisations are used.
__device__ void iir_line_lds(float* line, float* tmp, int L) {
    const float z = kPole;
    const int nt = blockDim.x;
    const int T = (L + nt - 1) / nt;
    const int s = threadIdx.x * T;
    const int e = min(s + T, L);
    if (s < L) {
        float c;
        int start;
        if (s - kHalo <= 0) {
            int hz = min(kHorizon, L);
            float sum = line[0];
            float zn = z;
            for (int n = 0; n < hz; ++n) { sum += zn * line[n]; zn *= z; }
            c = kLambda * sum;
            start = 0;
        } else {
            start = s - kHalo;
            c = kLambda * line[start];
        }
        if (start >= s) tmp[start] = c;
        for (int n = start + 1; n < e; ++n) {
            c = kLambda * line[n] + z * c;
            if (n >= s) tmp[n] = c;
        }
    }
    __syncthreads();
    if (s < L) {
        int end = min(e - 1 + kHalo, L - 1);
        float c = (end == L - 1) ? (z / (z - 1.0f)) * tmp[L - 1] : -z * tmp[end];
        if (end < e) line[end] = c;
        for (int n = end - 1; n >= s; --n) {
            c = z * (c - tmp[n]);
            if (n < e) line[n] = c;
        }
    }
    __syncthreads();
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
    for (int x = threadIdx.x; x < N; x += blockDim.x) {
        int xs = x - o;
        line[x] = (rowin && xs >= 0 && xs < N0) ? src[xs] : 0.f;
    }
    __syncthreads();
    if (cubic) iir_line_lds(line, tmp, N);
    float* dst = fc + ((size_t)z * N + yi) * N;
    for (int x = threadIdx.x; x < N; x += blockDim.x) dst[x] = line[x];
}

// adj input: copy3Dshifted (lpRgpu.cu:201) + stepangles (:202) + padker (:203) +
// applyFilter (:333-382) + s-direction prefilter.  The reference's FBP filter is a
// circular convolution of the 4N zero-padded row with a real even kernel; only lags
// |l| < N reach the N outputs, so it is evaluated exactly as a length-2N circular
// convolution with the kernel restricted to those lags (spectrum H2, built on host).
template <int EPT>
__global__ void k_sino_rows(const float* __restrict__ R, float* __restrict__ Rc, SinoPrep sp, int cubic) {
    extern __shared__ float2 smc[];
    const int N = sp.N;
    const int ai = blockIdx.x, z = blockIdx.y;
    float* dst = Rc + ((size_t)z * sp.P + ai) * N;
    if (ai % sp.os != 0) {
        for (int x = threadIdx.x; x < N; x += blockDim.x) dst[x] = 0.f;
        return;
    }
    const int a = ai / sp.os;
    const float* src = R + ((size_t)z * sp.P0 + a) * sp.N0;
    float* buf = reinterpret_cast<float*>(smc);
    float* tmp = buf + 2 * N;
    for (int x = threadIdx.x; x < N; x += blockDim.x) {
        int xs = min(max(x - sp.lb, 0), sp.N0 - 1);
        buf[x] = src[xs];
    }
    if (sp.H2) {
        for (int x = N + threadIdx.x; x < 2 * N; x += blockDim.x) buf[x] = 0.f;
        __syncthreads();
        fft_lds<EPT>(smc, N, 1, sp.planN, sp.twN, -1);
        r2c_post(smc, N, 1, N, sp.tw2N);
        for (int k = threadIdx.x; k < N; k += blockDim.x) {
            float2 v = smc[k];
            if (k == 0) smc[0] = make_float2(v.x * sp.H2[0], v.y * sp.H2[N]);
            else smc[k] = make_float2(v.x * sp.H2[k], v.y * sp.H2[k]);
        }
        __syncthreads();
        c2r_pre(smc, N, 1, N, sp.tw2N);
        fft_lds<EPT>(smc, N, 1, sp.planN, sp.twN, +1);
    }
    __syncthreads();
    if (cubic) iir_line_lds(buf, tmp, N);
    for (int x = threadIdx.x; x < N; x += blockDim.x) dst[x] = buf[x];
}

// y-direction prefilter (SamplesToCoefficients2DY after the transposes of
// prefilter2D, lpRgpu.cu:229-239): one thread per column, T rows per chunk with a
// kHalo-row halo on each side held in registers.
// Out of place: chunks read halos that neighbouring chunks write.
template <int T>
__global__ void k_col_iir(const float* __restrict__ in, float* __restrict__ outp, int L, int W) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int r0 = blockIdx.y * T;
    const int z = blockIdx.z;
    if (x >= W) return;
    constexpr int K = kHalo;
    constexpr int NV = T + 2 * K;
    const float* col = in + (size_t)z * L * W + x;
    float* ocol = outp + (size_t)z * L * W + x;
    float v[NV];
#pragma unroll
    for (int n = 0; n < NV; ++n) {
        int row = r0 - K + n;
        v[n] = (row >= 0 && row < L) ? col[(size_t)row * W] : 0.f;
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
__global__ void k_rows_fwd(RowPass rp, const float* __restrict__ F, const int2* __restrict__ range, Src src,
                           const int* __restrict__ xrow, const float4* __restrict__ xent, int cubic, int ldm) {
    extern __shared__ float2 smc[];
    const int Nt = rp.Ntheta, h = rp.h, R = rp.R, Nr = rp.Nrho;
    const int row0 = blockIdx.x * R, span = blockIdx.y;
    const int z0 = blockIdx.z * rp.zgroup;
    const int z1 = min(z0 + rp.zgroup, rp.nz);
    const int2* rng = range + (size_t)span * Nr + row0;
    for (int z = z0; z < z1; ++z) {
        const float* Fz = F + (((size_t)span * rp.nz + z) * Nr + row0) * (size_t)Nt;
        for (int p = threadIdx.x; p < R * Nt; p += blockDim.x) {
            int r = p / Nt, th = p - r * Nt;
            int2 lr = rng[r];
            reinterpret_cast<float*>(smc + r * ldm)[th] = (th >= lr.x && th < lr.y) ? Fz[p] : 0.f;
        }
        if (xrow) {
            __syncthreads();
            const float* T = src.p + (size_t)z * src.zstride;
            for (int r = 0; r < R; ++r) {
                int lo = xrow[span * (Nr + 1) + row0 + r], hi = xrow[span * (Nr + 1) + row0 + r + 1];
                for (int e = lo + threadIdx.x; e < hi; e += blockDim.x) {
                    float4 q = xent[e];
                    float v = sample(T, src.W, src.H, q.y, q.z, cubic);
                    atomicAdd(reinterpret_cast<float*>(smc + r * ldm) + (int)q.x, v);
                }
            }
        }
        __syncthreads();
        fft_lds<EPT>(smc, ldm, R, rp.planh, rp.twh, -1);
        r2c_post(smc, ldm, R, h, rp.twN);
        float2* out = rp.S + ((size_t)span * rp.nz + z) * h * (size_t)Nr + row0;
        for (int q = threadIdx.x; q < R * h; q += blockDim.x) {
            int f = q / R, r = q - f * R;
            out[(size_t)f * Nr + r] = smc[r * ldm + f];
        }
        __syncthreads();
    }
}

// LDS-tiled gather (the interp kernel of main_kernels.cuh:64-84 for all four
// grids): the block stages the source box of its output tile with coalesced
// row loads, then evaluates every sample of the tile from LDS.  A block owns
// `group` tiles that share one output tile (the three spans of the
// LP -> Cartesian gather, lpRgpu.cu:328, summed in span order).
__global__ void __launch_bounds__(kGatherThreads) k_gather(Gather g, const GTile* __restrict__ tiles) {
    extern __shared__ float box[];
    constexpr int PPT = kGatherPPT;
    constexpr int MAXG = 3;
    const int nt = blockDim.x;
    const GTile* tb = tiles + (size_t)blockIdx.x * g.group;
    const GTile t0 = tb[0];
    const int npts = t0.nr * t0.nc;
    const int z0 = blockIdx.y * g.zgroup, z1 = min(z0 + g.zgroup, g.nz);
    float2 cc[MAXG][PPT];
#pragma unroll
    for (int gi = 0; gi < MAXG; ++gi) {
        if (gi >= g.group) break;
        const GTile t = tb[gi];
#pragma unroll
        for (int e = 0; e < PPT; ++e) {
            int p = threadIdx.x + e * nt;
            cc[gi][e] = make_float2(__int_as_float(0x7fc00000), 0.f);
            if (p < npts) {
                int r = p / t.nc, c = p - r * t.nc;
                cc[gi][e] = g.coord[(size_t)t.span * g.c_span + (size_t)(t.r0 + r + g.c_ro) * g.c_row + t.c0 + c + g.c_co];
            }
        }
    }
    for (int z = z0; z < z1; ++z) {
        float acc[PPT];
#pragma unroll
        for (int e = 0; e < PPT; ++e) acc[e] = 0.f;
#pragma unroll
        for (int gi = 0; gi < MAXG; ++gi) {
            if (gi >= g.group) break;
            const GTile t = tb[gi];
            if (t.bw > 0) {
                // flattened box, 8 independent loads per thread in flight;
                // (i + 0.5) / bw in float is exact enough for i, bw < 2^20
                const float* S = g.src + (size_t)z * g.src_z + (size_t)t.span * g.src_span;
                const int nbox = t.bw * t.bh;
                const float inv = 1.0f / (float)t.bw;
                constexpr int UL = 8;
                for (int i0 = threadIdx.x; i0 < nbox; i0 += UL * nt) {
                    float v[UL];
#pragma unroll
                    for (int u = 0; u < UL; ++u) {
                        int i = i0 + u * nt;
                        v[u] = 0.f;
                        if (i < nbox) {
                            int yy = (int)(((float)i + 0.5f) * inv);
                            int xx = i - yy * t.bw;
                            v[u] = S[(size_t)wrapi(t.y0 + yy, g.H) * g.W + wrapi(t.x0 + xx, g.W)];
                        }
                    }
#pragma unroll
                    for (int u = 0; u < UL; ++u) {
                        int i = i0 + u * nt;
                        if (i < nbox) box[i] = v[u];
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (int e = 0; e < PPT; ++e) {
                float2 c = cc[gi][e];
                if (c.x == c.x) {
                    float fx = floorf(c.x), fy = floorf(c.y);
                    float ax = c.x - fx, ay = c.y - fy;
                    int lx = min(max((int)fx - 1 - t.x0, 0), t.bw - 4);
                    int ly = min(max((int)fy - 1 - t.y0, 0), t.bh - 4);
                    const float* q = box + ly * t.bw + lx;
                    float v;
                    if (g.cubic) {
                        float wx[4], wy[4];
                        bspline_w(ax, wx);
                        bspline_w(ay, wy);
                        v = 0.f;
#pragma unroll
                        for (int b = 0; b < 4; ++b) {
                            const float* qr = q + b * t.bw;
                            v += wy[b] * (wx[0] * qr[0] + wx[1] * qr[1] + wx[2] * qr[2] + wx[3] * qr[3]);
                        }
                    } else {
                        const float* q1 = q + t.bw + 1;
                        const float* q2 = q1 + t.bw;
                        v = (1.f - ay) * ((1.f - ax) * q1[0] + ax * q1[1]) + ay * ((1.f - ax) * q2[0] + ax * q2[1]);
                    }
                    acc[e] += v;
                }
            }
            __syncthreads();
        }
        float* O = g.out + (size_t)z * g.o_z + (size_t)t0.span * g.o_span;
#pragma unroll
        for (int e = 0; e < PPT; ++e) {
            int p = threadIdx.x + e * nt;
            if (p < npts) {
                int r = p / t0.nc, c = p - r * t0.nc;
                float v = acc[e];
                if (g.scale) v *= g.scale[(size_t)(t0.r0 + r) * g.s_row + t0.c0 + c];
                O[(size_t)(t0.r0 + r) * g.o_row + t0.c0 + c] = v;
            }
        }
    }
}

// Row pass, inverse direction: C2R along theta for R rows, writing G.
template <int EPT>
__global__ void k_rows_inv(RowPass rp, float* __restrict__ G, int ldm) {
    extern __shared__ float2 smc[];
    const int Nt = rp.Ntheta, h = rp.h, R = rp.R, Nr = rp.Nrho;
    const int row0 = blockIdx.x * R, span = blockIdx.y;
    const int z0 = blockIdx.z * rp.zgroup;
    const int z1 = min(z0 + rp.zgroup, rp.nz);
    for (int z = z0; z < z1; ++z) {
        const float2* in = rp.S + ((size_t)span * rp.nz + z) * h * (size_t)Nr + row0;
        for (int q = threadIdx.x; q < R * h; q += blockDim.x) {
            int f = q / R, r = q - f * R;
            smc[r * ldm + f] = in[(size_t)f * Nr + r];
        }
        __syncthreads();
        c2r_pre(smc, ldm, R, h, rp.twN);
        fft_lds<EPT>(smc, ldm, R, rp.planh, rp.twh, +1);
        float* out = G + (((size_t)span * rp.nz + z) * Nr + row0) * (size_t)Nt;
        for (int p = threadIdx.x; p < R * Nt; p += blockDim.x) {
            int r = p / Nt, th = p - r * Nt;
            out[p] = reinterpret_cast<const float*>(smc + r * ldm)[th];
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
        for (int i = threadIdx.x; i < M; i += blockDim.x) smc[i] = col[i];
        __syncthreads();
        fft_lds<EPT>(smc, M, 1, planM, twM, -1);
        if (f == 0) {
            for (int k = threadIdx.x; k <= M / 2; k += blockDim.x) {
                int km = (M - k) % M;
                float2 A = smc[k], Bc = cconj(smc[km]);
                float2 D = make_float2(0.5f * (A.x + Bc.x), 0.5f * (A.y + Bc.y));
                float2 E = make_float2(0.5f * (A.x - Bc.x), 0.5f * (A.y - Bc.y));
                float2 Q = make_float2(E.y, -E.x);  // E / i
                float2 yk = cadd(cmul(D, fz[k]), rot90(cmul(Q, fzh[k]), 1));
                float2 ykm = cadd(cmul(cconj(D), fz[km]), rot90(cmul(cconj(Q), fzh[km]), 1));
                smc[k] = yk;
                if (km != k) smc[km] = ykm;
            }
        } else {
            for (int i = threadIdx.x; i < M; i += blockDim.x) smc[i] = cmul(smc[i], fz[i]);
        }
        __syncthreads();
        fft_lds<EPT>(smc, M, 1, planM, twM, +1);
        for (int i = threadIdx.x; i < M; i += blockDim.x) col[i] = smc[i];
        __syncthreads();
    }
}

// fwd output: interp LP -> polar (lpRgpu.cu:284) for the kept angles only, fused
// with copy3Dstep (:191) and the detector crop copy3Dshifted (:192).
__global__ void k_polar_out(const float* __restrict__ G, float* __restrict__ R, int nz, int Nt, int Nr, int N,
                            int N0, int P0, int os, int lb, const float2* __restrict__ p2lp,
                            const unsigned char* __restrict__ span, int cubic, int zgroup) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int a = blockIdx.y;
    if (x >= N0) return;
    const int ai = a * os;
    const int k = span[ai];
    const float2 c = p2lp[(size_t)ai * N + x + lb];
    const int z0 = blockIdx.z * zgroup, z1 = min(z0 + zgroup, nz);
    for (int z = z0; z < z1; ++z) {
        const float* T = G + ((size_t)k * nz + z) * (size_t)Nr * Nt;
        R[((size_t)z * P0 + a) * N0 + x] = sample(T, Nt, Nr, c.x, c.y, cubic);
    }
}

// adj output: interp LP -> Cartesian accumulated over the spans (lpRgpu.cu:328),
// fused with the centre crop (:206).
__global__ void k_cart_out(const float* __restrict__ G, float* __restrict__ f, int nz, int Nt, int Nr, int N, int N0,
                           const float2* __restrict__ cmapC, int cubic, int zgroup) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= N0) return;
    const int o = N / 2 - N0 / 2;
    const size_t p = (size_t)(y + o) * N + x + o;
    const size_t NN = (size_t)N * N;
    float2 c[3];
    for (int k = 0; k < 3; ++k) c[k] = cmapC[k * NN + p];
    const int z0 = blockIdx.z * zgroup, z1 = min(z0 + zgroup, nz);
    const size_t grid = (size_t)Nr * Nt;
    for (int z = z0; z < z1; ++z) {
        float acc = 0.f;
        for (int k = 0; k < 3; ++k)
            if (c[k].x == c[k].x) acc += sample(G + ((size_t)k * nz + z) * grid, Nt, Nr, c[k].x, c[k].y, cubic);
        f[((size_t)z * N0 + y) * N0 + x] = acc;
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

size_t rows_lds_bytes(int ldm, int R) { return (size_t)R * ldm * sizeof(float2); }
size_t cols_lds_bytes(int M) { return (size_t)M * sizeof(float2); }
size_t sino_lds_bytes(int N) { return (size_t)(3 * N) * sizeof(float); }

#define LP_DISPATCH_EPT(ept, KER, grid, block, lds, stream, ...)                                   \
    do {                                                                                          \
        switch (ept) {                                                                            \
            case 4: hipLaunchKernelGGL(KER<4>, grid, block, lds, stream, __VA_ARGS__); break;     \
            case 8: hipLaunchKernelGGL(KER<8>, grid, block, lds, stream, __VA_ARGS__); break;     \
            case 16: hipLaunchKernelGGL(KER<16>, grid, block, lds, stream, __VA_ARGS__); break;   \
            default: hipLaunchKernelGGL(KER<32>, grid, block, lds, stream, __VA_ARGS__); break;   \
        }                                                                                         \
    } while (0)

static int threads_for(long long points) {
    int nt = 256;
    while (points > 32LL * nt && nt < 1024) nt *= 2;
    return nt;
}

void launch_image_rows(const float* f, float* fc, int nz, int N0, int N, bool cubic, hipStream_t s) {
    int nt = 256;
    hipLaunchKernelGGL(k_image_rows, dim3(N, nz), dim3(nt), 2 * N * sizeof(float), s, f, fc, N0, N,
                       N / 2 - N0 / 2, (int)cubic);
}

void launch_sino_rows(const float* R, float* Rc, int nz, const SinoPrep& sp, bool cubic, hipStream_t s) {
    int nt = threads_for(sp.N);
    int ept = ept_for(sp.N, nt);
    LP_DISPATCH_EPT(ept, k_sino_rows, dim3(sp.P, nz), dim3(nt), sino_lds_bytes(sp.N), s, R, Rc, sp, (int)cubic);
}

void launch_col_iir(const float* in, float* out, int nz, int L, int W, int chunk, hipStream_t s) {
    dim3 block(256);
    switch (chunk) {
        case 32: hipLaunchKernelGGL(k_col_iir<32>, dim3((W + 255) / 256, (L + 31) / 32, nz), block, 0, s, in, out, L, W); break;
        case 128: hipLaunchKernelGGL(k_col_iir<128>, dim3((W + 255) / 256, (L + 127) / 128, nz), block, 0, s, in, out, L, W); break;
        default: hipLaunchKernelGGL(k_col_iir<64>, dim3((W + 255) / 256, (L + 63) / 64, nz), block, 0, s, in, out, L, W); break;
    }
}

void launch_gather(const Gather& g, const GTile* tiles, int ntiles, hipStream_t s) {
    if (ntiles <= 0) return;
    dim3 grid(ntiles / g.group, (g.nz + g.zgroup - 1) / g.zgroup);
    hipLaunchKernelGGL(k_gather, grid, dim3(kGatherThreads), (size_t)g.box_floats * sizeof(float), s, g, tiles);
}

void launch_rows_fwd(const RowPass& rp, const float* F, const int2* range, const Src& src, const int* xrow,
                     const float4* xent, bool cubic, hipStream_t s) {
    int nt = threads_for((long long)rp.R * rp.h);
    int ept = ept_for((long long)rp.R * rp.h, nt);
    int ldm = rp.ldm;
    dim3 grid(rp.Nrho / rp.R, 3, (rp.nz + rp.zgroup - 1) / rp.zgroup);
    LP_DISPATCH_EPT(ept, k_rows_fwd, grid, dim3(nt), rows_lds_bytes(ldm, rp.R), s, rp, F, range, src, xrow, xent,
                    (int)cubic, ldm);
}

void launch_rows_inv(const RowPass& rp, float* G, hipStream_t s) {
    int nt = threads_for((long long)rp.R * rp.h);
    int ept = ept_for((long long)rp.R * rp.h, nt);
    int ldm = rp.ldm;
    dim3 grid(rp.Nrho / rp.R, 3, (rp.nz + rp.zgroup - 1) / rp.zgroup);
    LP_DISPATCH_EPT(ept, k_rows_inv, grid, dim3(nt), rows_lds_bytes(ldm, rp.R), s, rp, G, ldm);
}

void launch_cols(const RowPass& rp, const FftPlan& planM, const float2* twM, const float2* fzT, hipStream_t s) {
    int nt = threads_for(rp.Nrho);
    int ept = ept_for(rp.Nrho, nt);
    dim3 grid(rp.h, 3, (rp.nz + rp.zgroup - 1) / rp.zgroup);
    LP_DISPATCH_EPT(ept, k_cols, grid, dim3(nt), cols_lds_bytes(rp.Nrho), s, rp, planM, twM, fzT);
}

void launch_polar_out(const float* G, float* R, int nz, int Ntheta, int Nrho, int N, int N0, int P0, int os, int lb,
                      const float2* p2lp, const unsigned char* span, bool cubic, hipStream_t s) {
    dim3 grid((N0 + 255) / 256, P0, nz);
    hipLaunchKernelGGL(k_polar_out, grid, dim3(256), 0, s, G, R, nz, Ntheta, Nrho, N, N0, P0, os, lb, p2lp, span,
                       (int)cubic, 1);
}

void launch_cart_out(const float* G, float* f, int nz, int Ntheta, int Nrho, int N, int N0, const float2* cmapC,
                     bool cubic, hipStream_t s) {
    dim3 grid((N0 + 255) / 256, N0, nz);
    hipLaunchKernelGGL(k_cart_out, grid, dim3(256), 0, s, G, f, nz, Ntheta, Nrho, N, N0, cmapC, (int)cubic, 1);
}

}  // namespace lpk
