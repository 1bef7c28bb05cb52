/*
 * lprec_amd C ABI: log-polar Radon transform (fwd) and back-projection (adj)
 * on AMD Instinct MI355X (gfx950).
 *
 * Plain C types only.  Device buffers are raw pointers to float32 device
 * memory; `stream` is a hipStream_t passed as void* (NULL = the legacy
 * default stream).  Every call returns LP_OK (0) or a negative LP_ERR_*
 * code; lp_last_error() / lp_pre_error() give the message.  No call ever
 * resets the device.
 *
 * Each entry point below cites the reference interface it replaces
 * (paths relative to the lprec source tree).
 */
#ifndef LPREC_AMD_H
#define LPREC_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LP_OK 0
#define LP_ERR_ARG (-1)      /* invalid argument / shape / state            */
#define LP_ERR_ALLOC (-2)    /* host or device allocation failed            */
#define LP_ERR_DEVICE (-3)   /* HIP runtime or kernel launch error          */
#define LP_ERR_UNSUPPORTED (-4)

typedef struct lp_pre lp_pre;
typedef struct lp_handle lp_handle;

/* ---- host precompute (no GPU needed) -------------------------------------
 * Replaces lpTransform.precompute (src/lprec/lpTransform.py:19-31), i.e.
 * initsgl.create_gl (initsgl.py:25-57), initsfwd.create_fwd
 * (initsfwd.py:15-66) and initsadj.create_adj (initsadj.py:19-116).
 * `cor` is the float centre of rotation (used as float to size the padded
 * grid, truncated to int for the detector shift, as the reference does).
 * `filter` is one of none/ramp/shepp-logan/cosine/cosine2/hamming/hann/parzen
 * (case-insensitive; NULL = none).  `want_fwd` = the reference's flg. */
int lp_pre_create(int N0, int Nproj, int Nslices, double cor, int cubic, const char* filter, int want_fwd,
                  lp_pre** out);
/* sizes[0..4] = element counts of glpars, fwdparsi, fwdparamsf, adjparsi,
 * adjparamsf (the wire format of src/cuda/params.cu:40-145). */
int lp_pre_sizes(const lp_pre* pre, int64_t* sizes);
/* Copies the five arrays out (any pointer may be NULL to skip it). */
int lp_pre_copy(const lp_pre* pre, float* glpars, int32_t* fwdparsi, float* fwdparamsf, int32_t* adjparsi,
                float* adjparamsf);
void lp_pre_destroy(lp_pre* pre);
/* Message of the last failed lp_pre_* / lp_create call on this thread. */
const char* lp_pre_error(void);

/* ---- operator handle ------------------------------------------------------
 * lp_create replaces lpRgpu::lpRgpu(size_t params, int gpu)
 * (src/cuda/lpRgpu.cu:7-39, SWIG src/cuda/lpRgpu.i:16); the host array is
 * copied, device scratch is sized by Nslices (glpars[6]). */
int lp_create(const float* glpars, int device, lp_handle** out);
/* Replaces lpRgpu::initFwd (lpRgpu.cu:61-85; params.cu:57-99). */
int lp_init_fwd(lp_handle* h, const int32_t* fwdparsi, const float* fwdparamsf);
/* Replaces lpRgpu::initAdj (lpRgpu.cu:98-133; params.cu:101-145). */
int lp_init_adj(lp_handle* h, const int32_t* adjparsi, const float* adjparamsf);
/* Replaces lpRgpu::execFwdManyPtr(size_t Rptr, size_t fptr, int Nslices0, int gpu)
 * (lpRgpu.cu:184-193): f [ns0][N0][N0] -> R [ns0][Nproj0][N0], device memory.
 * Requires 1 <= ns0 <= Nslices. */
int lp_fwd(lp_handle* h, float* R, const float* f, int ns0, void* stream);
/* lp_fwd with no allocation of any kind: no reference counterpart.  For
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same operator and result as
 * lp_fwd, but the call's device scratch (prefiltered image, log-polar grids,
 * half spectra: 52 N^2 bytes per slice) comes from one static 64 MiB array
 * per GPU instead of the handle's buffers; larger calls (e.g. one slice at
 * N > 1024) return LP_ERR_ARG.  Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (other streams, other handles) race. */
int lp_fwd_noalloc(lp_handle* h, float* R, const float* f, int ns0, void* stream);
/* Replaces lpRgpu::execAdjManyPtr(size_t fptr, size_t Rptr, int Nslices0, int gpu)
 * (lpRgpu.cu:196-207): R [ns0][Nproj0][N0] -> f [ns0][N0][N0]; applies the FBP
 * filter when the adjoint tables carry one. */
int lp_adj(lp_handle* h, float* f, const float* R, int ns0, void* stream);
/* Exact transposes (no reference counterpart: the reference's adj is an
 * independently derived operator, initsadj.py:136 vs initsfwd.py:86, and is
 * only approximately the transpose of fwd).  lp_fwd_T computes f = F^T R for
 * the operator F of lp_fwd (R [ns0][Nproj0][N0] -> f [ns0][N0][N0]);
 * lp_adj_T computes R = A^T f for the operator A of lp_adj, including its
 * FBP filter (f [ns0][N0][N0] -> R [ns0][Nproj0][N0]).  They serve gradient
 * back-propagation and the dot-product test; they overwrite the handle's
 * scratch like lp_fwd / lp_adj. */
int lp_fwd_T(lp_handle* h, float* f, const float* R, int ns0, void* stream);
int lp_adj_T(lp_handle* h, float* R, const float* f, int ns0, void* stream);
/* Replaces lpRgpu::~lpRgpu (lpRgpu.cu:41-58).  Safe on NULL. */
void lp_destroy(lp_handle* h);
/* Message of the last failed call on this handle. */
const char* lp_last_error(const lp_handle* h);
/* Replaces lpRgpu::printCurrentGPUMemory (params.cu:147-155). */
int lp_mem_info(int device, size_t* free_bytes, size_t* total_bytes);

/* HBM streaming kernel (no reference counterpart; a measurement helper):
 * kind 0 copies n_floats floats src -> dst (16-byte non-temporal loads and
 * stores), 1 reads src (dst receives blocks * 256 float4 partial sums),
 * 2 writes dst.  n_floats must be a multiple of 4; blocks <= 0 picks 2048.
 * The benchmark reports its GB/s as the measured ceiling next to the
 * vendor peak. */
int lp_stream_bench(int device, float* dst, const float* src, size_t n_floats, int kind, int blocks, void* stream);

/* ---- introspection and tuning (no reference counterpart) ---------------- */
/* info[0..11] = N, N0, Ntheta, Nrho, Nproj(internal), Nproj0, os, Nslices,
 * cor(int), cubic, has_filter, device bytes held by the handle (clamped to
 * int64). */
int lp_info(const lp_handle* h, int64_t* info);
/* Kernel schedule knobs: key 0 = log-polar rows per row-FFT block (1..16),
 * key 1 = slices per block in the row/column passes (>=1; halved on launch
 * while a pass would have fewer than 2048 blocks), key 2 = rows per
 * prefilter column chunk (32, 64 or 128), key 3 = slices per block in the
 * tiled gathers (>=1), key 4 = FFT kernel family (1 = radix-16 register
 * kernels where the length is 16*k*16 with k in {1,2,4,8,16}; 2, the
 * default = the same with the inverse row pass at h = 1024 as a 32 x 32
 * register FFT with 32 lanes per row; 0 = the mixed-radix LDS kernels
 * everywhere), key 5 = LDS budget of a gather block
 * in floats (1024..32768; tiles are split until their source box fits; only
 * before lp_init_fwd / lp_init_adj), key 6 = forward operator's inverse row
 * pass and polar gather fused into one kernel (1) or run as two (0, the
 * default); the fused kernel applies with 16 rows per row-FFT block and the
 * radix-16 kernels, key 7 = lp_fwd_T's transpose of the polar gather as two
 * atomics-free gathers (1, the default, where the geometry allows it; reads
 * back 0 where it does not) or as the tiled scatter (0), key 8 = the tiled
 * scatters of lp_fwd_T / lp_adj_T flush their source boxes through a staging
 * area summed per destination tile (1, the default; no atomics into the
 * destination) or with global atomics (0), key 9 = LDS row stride of the
 * gather boxes, a 4-bit mask over the gathers image -> log-polar (bit 0),
 * log-polar -> polar (1), sinogram -> log-polar (2), log-polar -> Cartesian
 * (3): a set bit picks per box the least stride >= its width whose residue
 * keeps same-bank box slots at least ~5 slots apart, a clear bit uses
 * width | 1 (default 4; only before lp_init_fwd / lp_init_adj), key 10 =
 * row and column passes compiled at run time (hipRTC, at lp_init_fwd /
 * lp_init_adj) for transform lengths the radix-16 kernels do not cover (1,
 * the default; 0 = the generic mixed-radix kernels; reads back 1 only while
 * such passes are loaded), key 11 = threads per block of the compiled column
 * pass (256, 512 or 1024; 0 = automatic, the default; recompiles; reads back
 * the threads in use), key 12 = variant of the compiled passes, a 4-bit
 * mask (bit 0: the column pass loads the next slice's column while the
 * current one is transformed; bit 1: its index arithmetic is redone per slice
 * instead of held in registers; bit 2: at least 4 waves per SIMD; bit 3:
 * odd-radix work items of 4, 6 or 8 output pairs, whichever takes the fewest
 * rounds of the block's threads over the module's radix-R > 5 stages,
 * instead of 8; default 14; recompiles), key 13 = slices per block of
 * lp_fwd's log-polar -> polar gather (>= 1, default 4; each block reads its
 * tile's coordinates once), key 14 = lp_fwd's image -> log-polar gather
 * reads every tap from global memory, one thread per point (1), instead of
 * from LDS-staged tile boxes (0, the default): a measurement of the sampling
 * that a gather fused into the row pass would do.  Returns LP_ERR_ARG for bad
 * values. */
int lp_set_tuning(lp_handle* h, int key, int value);
int lp_get_tuning(const lp_handle* h, int key, int* value);

/* Stage timing (no reference counterpart; the reference only prints
 * wall-clock tic/toc, tests/timing.py).  While enabled, lp_fwd / lp_adj record
 * a HIP event on the call's stream between consecutive kernel launches.
 * lp_profile(h, 1) enables and clears, lp_profile(h, 0) disables.
 * lp_profile_read synchronises the recorded events and returns, for
 * which = 0 (fwd) or 1 (adj), the summed milliseconds of each stage over all
 * calls since enabling in ms[0..6], and the call count in *ncalls.
 * Stages: fwd = prep rows, column prefilter, image -> log-polar gather, row
 * FFT, column pass, inverse row FFT, polar gather; adj = sinogram rows, column
 * prefilter, sinogram -> log-polar gather, row FFT, column pass, inverse row
 * FFT, Cartesian gather. */
int lp_profile(lp_handle* h, int enable);
int lp_profile_read(lp_handle* h, int which, float* ms, int* ncalls);

/* Stage-wise testing (no reference counterpart).  lp_debug_stop(h, k) makes
 * the next lp_fwd / lp_adj calls return after stage k (1 = input rows,
 * 2 = column prefilter, 3 = gather into the log-polar grid, 4 = row FFT,
 * 5 = column pass, 6 = inverse row FFT; 0 = run everything).  lp_debug_copy
 * copies an internal device buffer to host memory (synchronising the device):
 * which = 0 prefiltered image [Nslices][N][N], 1 prefiltered sinogram
 * [Nslices][Nproj][N], 2 log-polar grids [3][ns0][Nrho][Ntheta] of the last
 * call, 3 half spectra [3][ns0][Ntheta/2][Nrho] (complex) of the last call,
 * 4 the row-sweep output of the last call (image or sinogram layout).
 * `bytes` must not exceed the buffer; returns LP_ERR_ARG otherwise. */
int lp_debug_stop(lp_handle* h, int stage);
int lp_debug_copy(lp_handle* h, int which, void* dst, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif /* LPREC_AMD_H */
