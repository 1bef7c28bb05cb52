/* TEST INFRASTRUCTURE: FP64 C restatement of the oracle's periodic B-spline
 * sampling (oracle/mirror.py interp_taps), used by the mirror at large sizes.
 *
 * Follows the interp kernel with cubicTex2D / linearTex2D of the reference
 * (src/include/main_kernels.cuh:9-84): the value at table coordinate (x, y)
 * of the periodic grid T[H][W] is
 *     sum_{a,b} w_a(fx) w_b(fy) T[(iy-1+b) mod H][(ix-1+a) mod W]
 * with ix = floor(x), fx = x - ix and the cubic B-spline weights of
 * main_kernels.cuh:9-19 (exact, not the texture unit's 8-bit fractions), or
 * the bilinear sum over (ix, ix+1) x (iy, iy+1) for linear sampling.
 * Wrap addressing: cudaAddressModeWrap (src/cuda/lpRgpu.cu:22-23).
 *
 * Pinned in tests/test_oracle.py against scipy.ndimage.map_coordinates
 * (order 3 / 1, mode 'grid-wrap', no prefilter), against the numpy
 * restatement and against a brute-force 16-tap sum.
 */
#include <math.h>
#include <stdint.h>

static inline int64_t wrap(int64_t i, int64_t n) {
    int64_t r = i % n;
    return r < 0 ? r + n : r;
}

/* T: [nz][H][W]; x, y: n coordinates; out: [nz][n] (overwritten). */
void oracle_interp(const double* T, int64_t nz, int64_t H, int64_t W, const double* x, const double* y, int64_t n,
                   int cubic, double* out) {
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; ++p) {
        const double fxf = floor(x[p]), fyf = floor(y[p]);
        const double fx = x[p] - fxf, fy = y[p] - fyf;
        const int64_t ix = (int64_t)fxf, iy = (int64_t)fyf;
        if (cubic) {
            double wx[4], wy[4];
            const double ox = 1.0 - fx, oy = 1.0 - fy;
            wx[0] = ox * ox * ox / 6.0;
            wx[1] = 2.0 / 3.0 - 0.5 * fx * fx * (2.0 - fx);
            wx[2] = 2.0 / 3.0 - 0.5 * ox * ox * (2.0 - ox);
            wx[3] = fx * fx * fx / 6.0;
            wy[0] = oy * oy * oy / 6.0;
            wy[1] = 2.0 / 3.0 - 0.5 * fy * fy * (2.0 - fy);
            wy[2] = 2.0 / 3.0 - 0.5 * oy * oy * (2.0 - oy);
            wy[3] = fy * fy * fy / 6.0;
            int64_t cx[4], ry[4];
            for (int a = 0; a < 4; ++a) cx[a] = wrap(ix - 1 + a, W);
            for (int b = 0; b < 4; ++b) ry[b] = wrap(iy - 1 + b, H) * W;
            for (int64_t z = 0; z < nz; ++z) {
                const double* Tz = T + z * H * W;
                double acc = 0.0;
                for (int b = 0; b < 4; ++b) {
                    double row = 0.0;
                    for (int a = 0; a < 4; ++a) row += wx[a] * Tz[ry[b] + cx[a]];
                    acc += wy[b] * row;
                }
                out[z * n + p] = acc;
            }
        } else {
            const int64_t c0 = wrap(ix, W), c1 = wrap(ix + 1, W);
            const int64_t r0 = wrap(iy, H) * W, r1 = wrap(iy + 1, H) * W;
            for (int64_t z = 0; z < nz; ++z) {
                const double* Tz = T + z * H * W;
                out[z * n + p] = (1.0 - fy) * ((1.0 - fx) * Tz[r0 + c0] + fx * Tz[r0 + c1]) +
                                 fy * ((1.0 - fx) * Tz[r1 + c0] + fx * Tz[r1 + c1]);
            }
        }
    }
}
