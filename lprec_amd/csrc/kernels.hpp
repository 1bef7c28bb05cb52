// Device kernels of the log-polar pipeline (see kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "fft_lds.hpp"

namespace lpk {

// pole and gain of the cubic B-spline prefilter (main_kernels.cuh:2, :124)
constexpr float kPole = -0.267949192431122706f;   // sqrt(3) - 2
constexpr float kLambda = 6.0f;                   // (1 - z)(1 - 1/z)
constexpr int kHorizon = 12;                      // causal init horizon (main_kernels.cuh:99)
constexpr int kHalo = 16;                         // |z|^17 < 2e-10: chunk halo of the IIR sweeps

struct Src {            // a stack of 2-D float grids read through bicubic / bilinear taps
    const float* p;     // [nz][H][W]
    int W, H;
    long long zstride;
};

// ---- prefilter + placement sweeps -----------------------------------------
// fwd: user f [nz][N0][N0] -> fc [nz][N][N], centred, x-direction prefilter
void launch_image_rows(const float* f, float* fc, int nz, int N0, int N, bool cubic, hipStream_t s);
// adj: user R [nz][P0][N0] -> Rc [nz][P][N]: shift by lb, zero-stuff by os,
// edge-replicate, optional FBP convolution, s-direction prefilter
struct SinoPrep {
    int N0, N, P0, P, os, lb;
    const float* H2;          // [N+1] real spectrum of the length-2N filter kernel (or null)
    FftPlan planN;            // complex FFT of length N (the 2N real transform)
    const float2* twN;        // W_N^j, j < N
    const float2* tw2N;       // W_{2N}^k, k <= N/2
};
void launch_sino_rows(const float* R, float* Rc, int nz, const SinoPrep& sp, bool cubic, hipStream_t s);
// y-direction prefilter of [nz][L][W] in place (columns), chunked with halo
void launch_col_iir(float* a, int nz, int L, int W, int chunk, hipStream_t s);

// ---- log-polar convolution passes ------------------------------------------
struct RowPass {
    int Ntheta, Nrho, h, R;   // R: log-polar rows per block
    int ldm;                  // LDS line stride (complex elements) of the row passes
    int nz, zgroup;           // slices, slices per block
    FftPlan planh;            // length h = Ntheta/2
    const float2* twh;        // W_h^j
    const float2* twN;        // W_Ntheta^k
    float2* S;                // [3][nz][h][Nrho]
};
// gather + e^rho + R2C along theta; cmap [3][Nrho][Ntheta] (x = NaN: empty)
void launch_rows_fwd(const RowPass& rp, const Src& src, const float2* cmap, const float* rowscale,
                     const int* xrow, const float4* xent, bool cubic, hipStream_t s);
// C2R along theta -> G [3][nz][Nrho][Ntheta]
void launch_rows_inv(const RowPass& rp, float* G, hipStream_t s);
// FFT along rho, x fZ, inverse, in place on S; fzT [h+1][Nrho] (scaled)
void launch_cols(const RowPass& rp, const FftPlan& planM, const float2* twM, const float2* fzT, hipStream_t s);

// ---- output gathers ---------------------------------------------------------
// fwd: G -> user R [nz][P0][N0]; p2lp [P][N] (theta idx, rho idx), span [P]
void launch_polar_out(const float* G, float* R, int nz, int Ntheta, int Nrho, int N, int N0, int P0, int os,
                      int lb, const float2* p2lp, const unsigned char* span, bool cubic, hipStream_t s);
// adj: sum over spans of G -> user f [nz][N0][N0]; cmapC [3][N][N]
void launch_cart_out(const float* G, float* f, int nz, int Ntheta, int Nrho, int N, int N0,
                     const float2* cmapC, bool cubic, hipStream_t s);

size_t rows_lds_bytes(int ldm, int R);
size_t cols_lds_bytes(int M);
size_t sino_lds_bytes(int N);
int ept_for(long long points, int nt);

}  // namespace lpk
