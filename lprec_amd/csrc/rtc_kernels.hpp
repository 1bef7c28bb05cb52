// Row and column passes of the log-polar convolution with every size a
// compile-time constant, instantiated at run time (hipRTC) for the lengths of
// a handle that the radix-16 register kernels do not cover.  Same data
// layout, block mapping and arithmetic as k_rows_fwd, k_cols and k_rows_inv
// of kernels.hip (the FFT stages of fft_ct.hpp follow fft_lds.hpp), so the
// results agree with those kernels; only the index arithmetic is resolved by
// the compiler.
#pragma once
#include "kernels.hpp"
#include "fft_ct.hpp"

namespace lpct {

template <class T>
__device__ __forceinline__ T ld_at(const T* base, unsigned off) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off);
}

__device__ __forceinline__ float2 fzval(const float2* __restrict__ fz, int i, int conj) {
    float2 v = fz[i];
    return conj ? make_float2(v.x, -v.y) : v;
}

// Global -> LDS staging of n elements with 8 independent loads in flight per
// thread (loads at clamped indices, stores guarded).
template <int NT, class Load, class Store>
__device__ __forceinline__ void stage_in(int n, Load load, Store store) {
    constexpr int U = 8;
    for (int i0 = threadIdx.x; i0 < n; i0 += U * NT) {
        float2 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = load(min(i0 + u * NT, n - 1));
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + u * NT;
            if (i < n) store(i, v[u]);
        }
    }
}

template <int... Rs>
constexpr int odd_slots() {
    return OddTables<Rs...>::size > 0 ? OddTables<Rs...>::size : 2;
}

// An opaque zero, produced inside a loop: table reads indexed with it are not
// hoisted out of the loop (the compiler would otherwise keep every stage's
// twiddles live in registers across the slices of a block).
__device__ __forceinline__ int opaque_zero() {
    int z;
    asm volatile("s_mov_b32 %0, 0" : "=s"(z));
    return z;
}

// Column pass (k_cols): FFT along rho of one theta-frequency column of S, the
// kernel-spectrum product (column 0 packs the theta-DC and Nyquist columns),
// inverse FFT.  Grid (h, 3, slice groups), NT threads, at least WPE waves per
// SIMD.  VAR bit 0: the next slice's column is loaded into registers while
// the current one is transformed; bit 1: the FFT's index arithmetic and
// twiddle reads are redone per slice (not hoisted into registers held across
// the slices of a block).
template <int M, int NT, int WPE, int VAR, int... Rs>
__global__ void __launch_bounds__(NT, WPE) k_cols_ct(lpk::RowPass rp, const float2* __restrict__ twM,
                                                     const float2* __restrict__ fzT) {
    constexpr bool PF = VAR & 1, OPAQUE = (VAR & 2) != 0;
    constexpr int PL = (M + NT - 1) / NT;   // column values per thread
    __shared__ float2 smc[lds_slots(M)], smd[lds_slots(M)];
    __shared__ __align__(16) float2 wq[odd_slots<Rs...>()];
    OddTables<Rs...>::template fill<M>(wq, twM, NT);
    const int h = rp.h;
    const int f = blockIdx.x, span = blockIdx.y;
    const int z0 = blockIdx.z * rp.zgroup;
    const int z1 = min(z0 + rp.zgroup, rp.nz);
    const float2* fz = fzT + (size_t)f * M;
    const float2* fzh = fzT + (size_t)h * M;
    const int2 ir = rp.in_rows ? rp.in_rows[span] : make_int2(0, M);
    const unsigned char* om = rp.out_rows ? rp.out_rows + (size_t)span * M : nullptr;
    auto column = [&](int z) { return rp.S + (((size_t)span * rp.nz + z) * h + f) * (size_t)M; };
    // column values of this thread (i = threadIdx.x + u NT), zero outside the
    // rows that carry data; loads at clamped indices, unconditional
    auto load_col = [&](int z, float2* r) {
        const float2* col = column(z);
#pragma unroll
        for (int u = 0; u < PL; ++u) {
            const int i = threadIdx.x + u * NT;
            const float2 v = col[min(i, M - 1)];
            r[u] = (i >= ir.x && i < ir.y) ? v : make_float2(0.f, 0.f);
        }
    };
    float2 nxt[PF ? PL : 1];
    if (PF && z0 < z1) load_col(z0, nxt);
    __syncthreads();
    for (int z = z0; z < z1; ++z) {
        float2* col = column(z);
        const int tid = OPAQUE ? (int)threadIdx.x + opaque_zero() : (int)threadIdx.x;
        if (PF) {
#pragma unroll
            for (int u = 0; u < PL; ++u) {
                const int i = threadIdx.x + u * NT;
                if (i < M) smc[lidx(i)] = nxt[u];
            }
            __syncthreads();
            if (z + 1 < z1) load_col(z + 1, nxt);
        } else {
            stage_in<NT>(
                M, [&](int i) { return (i >= ir.x && i < ir.y) ? col[i] : make_float2(0.f, 0.f); },
                [&](int i, float2 v) { smc[lidx(i)] = v; });
            __syncthreads();
        }
        float2* sp = fft<M, 1, M, NT>(smc, smd, twM, wq, -1, Radices<Rs...>{}, tid);
        if (f == 0) {
            for (int k = threadIdx.x; k <= M / 2; k += NT) {
                const int km = k == 0 ? 0 : M - k;
                const float2 A = sp[lidx(k)], Bc = cconj(sp[lidx(km)]);
                const float2 D = make_float2(0.5f * (A.x + Bc.x), 0.5f * (A.y + Bc.y));
                const float2 E = make_float2(0.5f * (A.x - Bc.x), 0.5f * (A.y - Bc.y));
                const float2 Q = make_float2(E.y, -E.x);  // E / i
                const float2 yk = cadd(cmul(D, fzval(fz, k, rp.conj)), rot90(cmul(Q, fzval(fzh, k, rp.conj)), 1));
                const float2 ykm =
                    cadd(cmul(cconj(D), fzval(fz, km, rp.conj)), rot90(cmul(cconj(Q), fzval(fzh, km, rp.conj)), 1));
                sp[lidx(k)] = yk;
                if (km != k) sp[lidx(km)] = ykm;
            }
        } else {
            for (int i = threadIdx.x; i < M; i += NT) sp[lidx(i)] = cmul(sp[lidx(i)], fzval(fz, i, rp.conj));
        }
        __syncthreads();
        const float2* xp = fft<M, 1, M, NT>(sp, sp == smc ? smd : smc, twM, wq, +1, Radices<Rs...>{}, tid);
        for (int i = threadIdx.x; i < M; i += NT)
            if (!om || om[i]) col[i] = xp[lidx(i)];
        __syncthreads();
    }
}

// Row pass, forward (k_rows_fwd): R log-polar rows of one span, real FFT of
// length 2H along theta (complex FFT of length H + post-processing), half
// spectra stored transposed into S.  Grid (Nrho / R, 3, slice groups).
template <int H, int R, int LDM, int NT, int WPE, int OPQ, int... Rs>
__global__ void __launch_bounds__(NT, WPE) k_rows_fwd_ct(lpk::RowPass rp, const float* __restrict__ F,
                                                    const int2* __restrict__ range) {
    const float2* __restrict__ twh = rp.twh;
    __shared__ float2 smc[lds_slots(R * LDM)], smd[lds_slots(R * LDM)];
    __shared__ __align__(16) float2 wq[odd_slots<Rs...>()];
    __shared__ int2 rng_s[R];
    OddTables<Rs...>::template fill<H>(wq, twh, NT);
    constexpr int Nt = 2 * H;
    const int Nr = rp.Nrho;
    const int row0 = blockIdx.x * R, span = blockIdx.y;
    const int z0 = blockIdx.z * rp.zgroup;
    const int z1 = min(z0 + rp.zgroup, rp.nz);
    if (threadIdx.x < R) rng_s[threadIdx.x] = range ? range[(size_t)span * Nr + row0 + threadIdx.x] : make_int2(0, Nt);
    __syncthreads();
    for (int z = z0; z < z1; ++z) {
        const float* Fz = F + (((size_t)span * rp.nz + z) * Nr + row0) * (size_t)Nt;
        stage_in<NT>(
            R * H,
            [&](int p) {
                // unconditional load at an index clamped into the row's range
                // (its lines are loaded anyway); pairs outside it are zeroed
                const int r = p / H, th = 2 * (p - r * H);
                const int2 lr = rng_s[r];
                const int tlo = min(max(lr.x & ~1, 0), Nt - 2), thi = max(tlo, min((lr.y - 1) & ~1, Nt - 2));
                float2 v = ld_at(reinterpret_cast<const float2*>(Fz), (unsigned)(r * Nt + min(max(th, tlo), thi)) * 4u);
                v.x = (th >= lr.x && th < lr.y) ? v.x : 0.f;
                v.y = (th + 1 >= lr.x && th + 1 < lr.y) ? v.y : 0.f;
                return v;
            },
            [&](int p, float2 v) {
                const int r = p / H, e = p - r * H;
                smc[lidx(r * LDM + e)] = v;
            });
        __syncthreads();
        float2* sp = fft<H, R, LDM, NT>(smc, smd, twh, wq, -1, Radices<Rs...>{},
                                        (int)threadIdx.x + (OPQ ? opaque_zero() : 0));
        r2c_post(sp, LDM, R, H, rp.twN);
        float2* out = rp.S + ((size_t)span * rp.nz + z) * H * (size_t)Nr + row0;
        for (int q = threadIdx.x; q < R * H; q += NT) {
            const int fr = q / R, r = q - fr * R;
            out[(size_t)fr * Nr + r] = sp[lidx(r * LDM + fr)];
        }
        __syncthreads();
    }
}

// Row pass, inverse (k_rows_inv): C2R along theta for R rows, writing the
// theta columns the output gather reads.
template <int H, int R, int LDM, int NT, int WPE, int OPQ, int... Rs>
__global__ void __launch_bounds__(NT, WPE) k_rows_inv_ct(lpk::RowPass rp, float* __restrict__ G,
                                                    const int2* __restrict__ wrange) {
    const float2* __restrict__ twh = rp.twh;
    __shared__ float2 smc[lds_slots(R * LDM)], smd[lds_slots(R * LDM)];
    __shared__ __align__(16) float2 wq[odd_slots<Rs...>()];
    OddTables<Rs...>::template fill<H>(wq, twh, NT);
    constexpr int Nt = 2 * H;
    const int Nr = rp.Nrho;
    const int row0 = blockIdx.x * R, span = blockIdx.y;
    const int z0 = blockIdx.z * rp.zgroup;
    const int z1 = min(z0 + rp.zgroup, rp.nz);
    __syncthreads();
    for (int z = z0; z < z1; ++z) {
        const float2* in = rp.S + ((size_t)span * rp.nz + z) * H * (size_t)Nr + row0;
        stage_in<NT>(
            R * H,
            [&](int q) {
                const int fr = q / R, r = q - fr * R;
                return in[(size_t)fr * Nr + r];
            },
            [&](int q, float2 v) {
                const int fr = q / R, r = q - fr * R;
                smc[lidx(r * LDM + fr)] = v;
            });
        __syncthreads();
        c2r_pre(smc, LDM, R, H, rp.twN);
        const float2* xp = fft<H, R, LDM, NT>(smc, smd, twh, wq, +1, Radices<Rs...>{},
                                              (int)threadIdx.x + (OPQ ? opaque_zero() : 0));
        float* out = G + (((size_t)span * rp.nz + z) * Nr + row0) * (size_t)Nt;
        for (int p = threadIdx.x; p < R * H; p += NT) {
            const int r = p / H, e = p - r * H;
            const int2 wr = wrange ? wrange[(size_t)span * Nr + row0 + r] : make_int2(0, Nt);
            if (2 * e + 1 >= wr.x && 2 * e < wr.y)
                *reinterpret_cast<float2*>(out + (size_t)r * Nt + 2 * e) = xp[lidx(r * LDM + e)];
        }
        __syncthreads();
    }
}

// Filtered sinogram rows (k_sino_rows_pair): two kept angles per block, the
// FBP filter as one complex FFT pair of length M = 2N of x0 + i x1 (the
// filter spectrum is real and even), then both lines' prefilters side by
// side.  The rows are read straight into the FFT buffer; the buffer the
// inverse FFT does not end in holds the two lines and the prefilter's
// scratch.  Grid ((P0 + 1) / 2, nz), NT threads.
template <int N, int NT, int WPE, int... Rs>
__global__ void __launch_bounds__(NT, WPE) k_sino_rows_pair_ct(const float* __restrict__ R, float* __restrict__ Rc,
                                                               lpk::SinoPrep sp, int cubic) {
    constexpr int M = 2 * N;
    __shared__ float2 sa[lds_slots(M)], sb[lds_slots(M)];
    __shared__ __align__(16) float2 wq[odd_slots<Rs...>()];
    OddTables<Rs...>::template fill<M>(wq, sp.tw2Nc, NT);
    const int a0 = 2 * blockIdx.x, z = blockIdx.y;
    const int nrow = min(2, sp.P0 - a0);
    for (int u = 0; u < (cubic ? 0 : nrow); ++u)
        for (int j = 1; j < sp.os; ++j) {
            float* dz = Rc + ((size_t)z * sp.P + (size_t)(a0 + u) * sp.os + j) * N;
            for (int x = threadIdx.x; x < N; x += NT) dz[x] = 0.f;
        }
    {
        const float* src0 = R + ((size_t)z * sp.P0 + a0) * sp.N0;
        const float* src1 = R + ((size_t)z * sp.P0 + min(a0 + 1, sp.P0 - 1)) * sp.N0;
        for (int x0 = threadIdx.x; x0 < M; x0 += 4 * NT) {
            float v0[4], v1[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int xs = min(max(x0 + u * NT - sp.lb, 0), sp.N0 - 1);
                v0[u] = src0[xs];
                v1[u] = src1[xs];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int x = x0 + u * NT;
                if (x < M) sa[lidx(x)] = x < N ? make_float2(v0[u], nrow > 1 ? v1[u] : 0.f) : make_float2(0.f, 0.f);
            }
        }
    }
    __syncthreads();
    const int tid = threadIdx.x;
    float2* xp = fft<M, 1, M, NT>(sa, sb, sp.tw2Nc, wq, -1, Radices<Rs...>{}, tid);
    for (int k = tid; k < M; k += NT) {
        const float hk = sp.H2[k <= N ? k : M - k];
        const float2 v = xp[lidx(k)];
        xp[lidx(k)] = make_float2(v.x * hk, v.y * hk);
    }
    __syncthreads();
    float2* yp = fft<M, 1, M, NT>(xp, xp == sa ? sb : sa, sp.tw2Nc, wq, +1, Radices<Rs...>{}, tid);
    float* line = reinterpret_cast<float*>(yp == sa ? sb : sa);   // [2][N] lines, then [2][N] scratch
    float* tmp = line + 2 * N;
    for (int e = tid; e < N; e += NT) {
        const float2 v = yp[lidx(e)];
        line[e] = v.x;
        line[N + e] = v.y;
    }
    __syncthreads();
    // the two lines' prefilters run side by side on the two halves of the block
    constexpr int half = NT / 2;
    const int u = tid < half ? 0 : 1;
    if (cubic) lpk::iir_line_lds(line + u * N, tmp + u * N, N, tid - u * half, half);
    for (int v = 0; v < nrow; ++v) {
        float* dst = Rc + ((size_t)z * sp.P + (size_t)(a0 + v) * sp.os) * N;
        for (int x = tid; x < N; x += NT) dst[x] = line[v * N + x];
    }
}

}  // namespace lpct
