// Device kernels of the log-polar pipeline (see kernels.hip).
#pragma once
#ifdef __HIPCC_RTC__
// compiled at run time (rtc_kernels.hpp): hipRTC supplies the device runtime;
// the launch declarations below only need the stream type
typedef struct ihipStream_t* hipStream_t;
#else
#include <hip/hip_runtime.h>
#endif

#include "fft_lds.hpp"

namespace lpk {

// pole and gain of the cubic B-spline prefilter (main_kernels.cuh:2, :124)
constexpr float kPole = -0.267949192431122706f;   // sqrt(3) - 2
constexpr float kLambda = 6.0f;                   // (1 - z)(1 - 1/z)
constexpr int kHorizon = 12;                      // causal init horizon (main_kernels.cuh:99)
constexpr int kHalo = 16;                         // |z|^17 < 2e-10: chunk halo of the IIR sweeps

// Cubic B-spline prefilter of one line held in LDS (ConvertToInterpolationCoefficients,
// main_kernels.cuh:97-142).  Each thread owns a chunk and restarts the causal and
// anticausal recursions kHalo samples outside it (truncation |z|^17 < 2e-10), except
// at the line ends where the reference's own initialisations are used.
// (tid, nt: this thread's index among the nt threads that share the line;
// every thread of the block calls it, since it synchronises the block)
__device__ inline void iir_line_lds(float* line, float* tmp, int L, int tid, int nt) {
    const float z = kPole;
    // odd chunk length: lanes start T floats apart, so an odd T keeps the
    // lanes of a wave on distinct banks
    const int T = ((L + nt - 1) / nt) | 1;
    const int s = tid * T;
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
    int Lh;                   // complex length of the filter transform (2 Lh >= 2N real)
    const float* H2;          // [Lh+1] real spectrum of the length-2Lh filter kernel (or null)
    FftPlan planN;            // complex FFT of length Lh (the 2Lh real transform)
    const float2* twN;        // W_Lh^j, j < Lh
    const float2* tw2N;       // W_{2Lh}^k, k <= Lh/2
    FftPlan plan2N;           // complex FFT of length 2N (two filtered lines packed as x0 + i x1)
    const float2* tw2Nc;      // W_{2N}^j, j < 2N
    // the filtered pair pass compiled at run time for this N (hipFunction_t of
    // rtc_kernels.hpp, rtc_pair_nt threads; null: the generic kernel)
    void* rtc_pair;
    int rtc_pair_nt;
};
void launch_sino_rows(const float* R, float* Rc, int nz, const SinoPrep& sp, bool cubic, hipStream_t s);
// y-direction prefilter of [nz][L][W] (columns), chunked with halo, in -> out
// (out of place: the chunks overlap through their halos); with stuff > 1 the
// input rows that are not multiples of stuff are zero and are not read
void launch_col_iir(const float* in, float* out, int nz, int L, int W, int chunk, int stuff, hipStream_t s);

// ---- log-polar convolution passes ------------------------------------------
struct RowPass {
    int Ntheta, Nrho, h, R;   // R: log-polar rows per block
    int ldm;                  // LDS line stride (complex elements) of the row passes
    int nz, zgroup;           // slices, slices per block
    FftPlan planh;            // length h = Ntheta/2
    const float2* twh;        // W_h^j
    const float2* twN;        // W_Ntheta^k
    float2* S;                // [3][nz][h][Nrho]
    int conj;                 // column pass multiplies by conj(fZ) (the transposed convolution)
    int fft16;                // use the radix-16 register kernels where the length allows
    // Rows of the log-polar grids that carry data (null = all rows):
    // in_rows[span] = [lo, hi), multiples of 128, outside which the gathered
    // grid is zero -- the row pass writes no half-spectrum rows there and the
    // column pass reads none (zeros); out_rows[span * Nrho + row] != 0 for the
    // rows the output gather reads -- the column pass stores only those and
    // the inverse row pass skips blocks with none.
    const int2* in_rows;
    const unsigned char* out_rows;
    // Row and column passes compiled at run time for this handle's lengths
    // (hipFunction_t of rtc_kernels.hpp; null: the mixed-radix kernels run):
    // rows_fwd / rows_inv with rtc_R rows and rtc_rows_nt threads per block,
    // cols with rtc_cols_nt threads.  Used where the radix-16 kernels do not
    // apply.
    void* rtc_rows_fwd;
    void* rtc_rows_inv;
    void* rtc_cols;
    int rtc_R, rtc_rows_nt, rtc_cols_nt;
};
// R2C along theta of the gathered log-polar grids F [3][nz][Nrho][Ntheta];
// range [3][Nrho]: [lo, hi) theta bounds of the non-empty points of each row
// (F is read only inside them, zero elsewhere; null = whole rows)
void launch_rows_fwd(const RowPass& rp, const float* F, const int2* range, hipStream_t s);
// F[span][z][row][theta] += sample(src_z, x, y) for the nx extra samples
// (theta, x, y, row) of xent (those sharing a log-polar point with another
// sample); xrow[span * (Nrho + 1)] is the first entry of each span
void launch_add_extras(float* F, int nz, int Ntheta, int Nrho, const Src& src, const int* xrow, const float4* xent,
                       int nx, bool cubic, hipStream_t s);
// C2R along theta -> G [3][nz][Nrho][Ntheta]; only the theta columns
// [lo, hi) of wrange [3][Nrho] are stored (null: whole rows)
void launch_rows_inv(const RowPass& rp, float* G, const int2* wrange, hipStream_t s);
// The inverse row pass fused with the log-polar -> polar gather (forward
// operator): bands of R - 3 base rows per block (R = 16), the polar samples
// of each (span, band) listed in band_beg order.
struct PolarOut {
    float* out;               // R [nz][Nproj0][N0]
    long long o_z;            // slice stride of out
    const float2* coord;      // per sample: (theta, rho) table coordinates
    const int* off;           // per sample: offset a * N0 + x in a slice of out
    const int* band_beg;      // [3][nband + 1] first sample of each (span, band)
    int nband;
    int cubic;
};
// false when the register kernels do not apply (the caller runs the unfused pair)
bool launch_rows_inv_polar(const RowPass& rp, const PolarOut& po, hipStream_t s);
// FFT along rho, x fZ, inverse, in place on S; fzT [h+1][Nrho] (scaled)
void launch_cols(const RowPass& rp, const FftPlan& planM, const float2* twM, const float2* fzT, hipStream_t s);

// ---- LDS-tiled gathers ---------------------------------------------------------
// One output tile (nr x nc points at (r0, c0)) and the source box [x0, x0+bw) x
// [y0, y0+bh) (unwrapped, periodic in the source) that holds every tap of its
// points for span `span`.  Built on the host from the coordinate tables; nc is
// a power of two (nc = 2^lnc <= kGatherThreads) and nr * nc <= kGatherThreads *
// kGatherPPT.
struct GTile {
    int r0, c0, nr, nc;
    int x0, y0, bw, bh;
    int span, lnc;
    // LDS row stride of the staged box (>= bw)
    int bs;
    // colour classes of the transposed (scatter) use: point (r, c) has colour
    // (r mod cr) * cc + (c mod cc), and points of one colour have disjoint
    // 4 x 4 tap footprints; cr = 0: no such colouring (atomic accumulation).
    // The tile's points with a sample, sorted by colour, are perm[pbase ..]
    // (r * nc + c), colour k at positions cls[cbase + k] .. cls[cbase + k + 1].
    int cr, cc;
    int pbase, cbase;
};
// out[z*o_z + span*o_span + r*o_row + c] = (sum over the `group` tiles of a
// block of) sample(src + z*src_z + span*src_span, coord[span*c_span +
// (r + c_ro)*c_row + c + c_co]) * scale[r*s_row + c]
struct Gather {
    const float* src;
    long long src_z, src_span;
    int W, H;
    const float2* coord;
    long long c_span;
    int c_row, c_ro, c_co;
    const float* scale;
    int s_row;
    float* out;
    long long o_z, o_span;
    int o_row;
    int nz, zgroup, group, cubic, box_floats;
};
constexpr int kGatherThreads = 256;
constexpr int kGatherPPT = 8;   // points per thread: tiles hold <= 2048 points
// slices per colour pass of the coloured scatter (k_scatter_col); 2 measured
// no faster (0.692 against 0.699 ms per call, N0 = 1024, 16 slices) for
// twice the LDS
constexpr int kScatterSlices = 1;
#ifndef LP_GATHER_PIPE
#define LP_GATHER_PIPE 2
#endif
constexpr int kGatherPipe = LP_GATHER_PIPE;   // points whose coordinates are in flight ahead of the sampling
// kind: 0 image->LP, 1 sinogram->LP, 2 LP->polar, 3 LP->Cartesian (names the kernel instance)
void launch_gather(const Gather& g, const GTile* tiles, int ntiles, int kind, hipStream_t s);
// The image -> log-polar gather with every tap read from global memory, one
// thread per point, only the theta range [3][Nrho] of each row (a measurement
// of the sampling a row pass fused with the gather would do)
void launch_gather_direct(const Gather& g, const int2* range, int Nr, int Nt, hipStream_t s);

// ---- transposes (exact adjoints of the stages above) -----------------------
// Transpose of a Gather: the value at every output point of every tile (times
// scale) is added into the taps of the gather's source grid, which is `dst`
// here (accumulated with atomics: dst must be cleared first).
struct Scatter {
    float* dst;
    long long dst_z, dst_span;
    int W, H;
    const float2* coord;
    long long c_span;
    int c_row, c_ro, c_co;
    const float* scale;
    int s_row;
    const float* val;
    long long v_z, v_span;
    int v_row;
    int nz, zgroup, group, cubic, box_floats;
    int tstride;   // set by launch_scatter: block order over the tile groups
    // staged flush (no atomics into dst): when stage is set, each block writes
    // its accumulated source box, bw x bh row-major, to
    // stage + z * stage_z + box_off[tile * group + gi]; launch_box_sum then
    // forms dst from the staged boxes
    float* stage;
    long long stage_z;
    const int* box_off;
    const unsigned short* perm;   // colour-sorted points of the tiles (GTile.pbase)
    const int* cls;               // colour class bounds (GTile.cbase)
};
// Destination of staged scatter boxes: dst tiles of 32 x 32 points, each the
// sum, in increasing box order, of the staged boxes that cover it (periodic
// placement); every dst point is written.
struct BoxSum {
    const float* stage;
    long long stage_z;
    const int* tile_beg;      // [ntiles + 1]: boxes of destination tile t in tile_box[tile_beg[t] ..)
    const int* tile_box;
    const int4* box_rect;     // per box: x0, y0 (unwrapped), bw, bh
    const int* box_off;       // per box: offset in a slice's staging
    float* dst;
    long long dst_z, dst_plane;
    int W, H, tiles_x, tiles_y, nplanes, nz;
};
void launch_box_sum(const BoxSum& b, hipStream_t s);
void launch_scatter(const Scatter& g, const GTile* tiles, int ntiles, int kind, hipStream_t s);

// Transpose of the log-polar -> polar gather as two gathers.  The polar map
// is separable: an angle's samples share one theta coordinate, and their rho
// coordinate grows with the detector cell.  Step A: u[z][k][a][r] = sum over
// the detector cells x of angle a whose rho taps include row r of the rho
// weight times R[z][a][x] (the cells form one run per (a, r)); step B:
// G[k][z][r][j] = sum over the angles of span k whose theta taps include
// column j of the theta weight times u[z][k][a][r].  Every output is summed
// in a fixed order by one thread: no atomics.
struct PolarT {
    const float* val;         // R [nz][P0][N0]
    long long v_z;
    const float2* coord;      // coordinate of (a, x): coord[a * c_row + c_co + x]
    int c_row, c_co;
    const float* cxa;         // theta coordinate of each kept angle
    float* u;                 // [nz][3][nA][Nr]
    int nA;                   // angles per span (padded)
    const int* abeg;          // [6] kept angles of span k: [abeg[k], abeg[3 + k])
    const int* xs;            // [P0][Q + 1]: first x with floor(rho) >= qlo + i (N0 if none)
    int qlo, Q;
    float* G;                 // [3][nz][Nr][Nt]
    const int* aj;            // [3][jn + 5]: first local angle of span k with floor(theta) >= jlo - 2 + i
    int jlo, jn;              // step B writes columns [jlo, jlo + jn)
    int Nr, Nt, N0, nz, cubic;
};
void launch_polar_transpose(const PolarT& p, hipStream_t s);
// transpose of the extra samples of launch_rows_fwd: F [3][nz][Nrho][Ntheta]
// values scattered into dst [nz][H][W] at the extras' coordinates
void launch_scatter_extra(const float* F, float* dst, int nz, int Ntheta, int Nrho, int W, int H,
                          const int* xrow, const float4* xent, bool cubic, hipStream_t s);
// transpose of launch_col_iir (same chunking)
void launch_col_iir_T(const float* in, float* out, int nz, int L, int W, int chunk, hipStream_t s);
// transpose of launch_image_rows: fc [nz][N][N] -> f [nz][N0][N0]
void launch_image_rows_T(const float* fc, float* f, int nz, int N0, int N, bool cubic, hipStream_t s);
// transpose of launch_sino_rows: Rc [nz][P][N] -> R [nz][P0][N0]
void launch_sino_rows_T(const float* Rc, float* R, int nz, const SinoPrep& sp, bool cubic, hipStream_t s);
// p[0..n) = 0 (a kernel, so that it may be captured in a graph)
void launch_zero(float* p, size_t n, hipStream_t s);

// HBM streaming kernels (kind 0 copy, 1 read, 2 write) over n4 float4 with
// `blocks` blocks of 256 lanes; for kind 1 dst must hold blocks * 256 float4.
int launch_stream(int kind, float* dst, const float* src, size_t n4, int blocks, hipStream_t s);

size_t rows_lds_bytes(int ldm, int R);
size_t cols_lds_bytes(int M);
size_t sino_lds_bytes(int Lh, int N);
int ept_for(long long points, int nt);

}  // namespace lpk
