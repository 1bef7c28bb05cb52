// Operator handle: owns the device tables and scratch of one geometry on one
// GPU and drives the kernels of kernels.hip.  Replaces the reference's class
// lpRgpu (src/include/lpRgpu.cuh:6-81, src/cuda/lpRgpu.cu, params.cu,
// gridStorage.cu) behind the C ABI of include/lprec_amd.h.
#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <dlfcn.h>
#include <hip/hiprtc.h>

#include "fft64.hpp"
#include "kernels.hpp"
#include "lp_internal.hpp"

namespace {

thread_local std::string g_create_err;

struct DeviceGuard {
    int prev = -1;
    bool ok = false;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        ok = hipSetDevice(dev) == hipSuccess;
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

FftPlan make_plan(int M) {
    FftPlan p;
    p.M = M;
    p.nst = 0;
    int n = M;
    std::vector<int> r;
    while (n % 8 == 0) { r.push_back(8); n /= 8; }
    while (n % 4 == 0) { r.push_back(4); n /= 4; }
    while (n % 2 == 0) { r.push_back(2); n /= 2; }
    for (int q : {3, 5}) while (n % q == 0) { r.push_back(q); n /= q; }
    std::vector<int> big;   // factors > 5 run first: their stage then needs no twiddles
    for (int q = 7; (long long)q * q <= n; q += 2) while (n % q == 0) { big.push_back(q); n /= q; }
    if (n > 1) big.push_back(n);
    r.insert(r.begin(), big.rbegin(), big.rend());
    if ((int)r.size() > LP_MAX_STAGES) throw std::runtime_error("FFT length has too many factors");
    for (size_t i = 0; i < r.size(); ++i) p.radix[i] = r[i];
    p.nst = (int)r.size();
    return p;
}

}  // namespace

struct lp_handle {
    int device = 0;
    int N = 0, N0 = 0, Nt = 0, Nr = 0, P = 0, P0 = 0, Ns = 0, cor = 0, os = 1, h = 0;
    bool cubic = true;
    bool has_fwd = false, has_adj = false, has_filter = false;
    // fuse_polar: off by default -- the fused kernel's 16-row windows start at
    // rows 13 b - 1, so each reads two 128-byte lines per spectrum column
    // where the separate pass reads one, and it measured no faster than the
    // separate pair (1.91 vs 1.90 ms per 32 slices at N = 2048)
    int R = 16, zgroup = 8, colchunk = 32, gzgroup = 4, fft16 = 2, fuse_polar = 0;
    int polar_t = 1;   // separable polar transpose where the geometry allows (else the tiled scatter)
    // slices per block of the log-polar -> polar gather: its coordinate table
    // ([P][N] float2) is read once per block.  8 measured 0.587 against 0.595
    // ms per 32 slices at N = 2048 but read more of G from HBM (2119 against
    // 2028 MB: the blocks of neighbouring tiles, which share G rows, run
    // further apart), so the default stays at 4
    int pzgroup = 4;
    // image -> log-polar gather from global memory, no LDS tiles (measurement
    // of the sampling a fused gather + row pass would do; off)
    int gather_direct = 0;
    // row / column passes compiled at run time for lengths the radix-16
    // kernels do not cover (rtc_build): setting, module, functions (rows
    // fwd, rows inv, cols) and their launch shapes
    int rtc_mode = 1;
    int rtc_cols_req = 0;   // threads of the compiled column pass (0: automatic)
    // variant of the compiled passes: k_cols_ct VAR (bits 0-1), at least 4
    // waves per SIMD (bit 2), odd-radix output pairs per item 4 instead of 8 (bit 3)
    int rtc_cols_var = 14;
    int rtc_pp = 0;   // output pairs per odd-radix work item of the loaded module
    bool rtc_tried = false;
    hipModule_t rtc_mod = nullptr;
    void* rtc_fn[4] = {nullptr, nullptr, nullptr, nullptr};   // + the filtered sinogram pair pass
    int rtc_R = 0, rtc_rows_nt = 0, rtc_cols_nt = 0, rtc_pair_nt = 0;
    size_t lds_limit = 65536;
    std::string err;
    std::vector<void*> allocs;
    size_t bytes = 0;

    float* d_erho = nullptr;
    float* d_G = nullptr;
    float* d_tmp = nullptr;   // row-sweep output feeding the column sweep (aliases G when it fits)
    float2* d_S = nullptr;
    std::map<int, float2*> tw;
    FftPlan plan_h{}, plan_N{}, plan_M{}, plan_F{}, plan_2N{};
    int Lf = 0;   // complex length of the FBP filter transform

    float* d_fc = nullptr;
    float2* d_cmapF = nullptr;
    float2* d_p2lp = nullptr;
    unsigned char* d_spanrow = nullptr;
    float2* d_fzF = nullptr;
    int2* d_rangeF = nullptr;
    int2* d_rangeA = nullptr;
    int2* d_wrangeF = nullptr;   // theta columns of G each (span, row) the polar gather reads
    int2* d_wrangeA = nullptr;   // ... and the Cartesian gather reads
    // staged (atomics-free) flush of the tiled scatters of the transposes:
    // per tile list, each box's offset and rectangle and, per destination
    // tile of 32 x 32, the boxes that cover it
    struct Stage {
        int* off = nullptr;
        int4* rect = nullptr;
        int* tbeg = nullptr;
        int* tbox = nullptr;
        long long total = 0;   // floats per slice
        int tiles_x = 0, tiles_y = 0, nplanes = 1;
        bool ok = false;
    } stF, stA, stC;
    int staged = 1;   // use the staged flush where the half-spectrum buffer holds it
    // colour-sorted points of the tile lists (atomics-free scatters)
    unsigned short *d_permF = nullptr, *d_permP = nullptr, *d_permA = nullptr, *d_permC = nullptr;
    int *d_clsF = nullptr, *d_clsP = nullptr, *d_clsA = nullptr, *d_clsC = nullptr;
    // separable transpose of the polar gather (pt_ok false: the geometry does not allow it)
    int* d_pt_abeg = nullptr;
    float* d_pt_cxa = nullptr;
    int* d_pt_xs = nullptr;
    int* d_pt_aj = nullptr;
    int pt_nA = 0, pt_qlo = 0, pt_Q = 0, pt_jlo = 0, pt_jn = 0;
    bool pt_ok = false;
    // the fused inverse row pass + polar gather: samples per (span, band of 13 base rows)
    float2* d_pcoord = nullptr;
    int* d_poff = nullptr;
    int* d_pband = nullptr;
    int npband = 0;
    // rows of the log-polar grids the convolution passes may skip (null: none)
    int2* d_inrowsF = nullptr;            // [3] rows [lo, hi) where the gathered grid can be non-zero
    int2* d_inrowsA = nullptr;
    unsigned char* d_outrowsF = nullptr;  // [3][Nrho] rows the output gather reads
    unsigned char* d_outrowsA = nullptr;
    // LDS-tiled gathers: tile lists and the largest source box of each
    lpk::GTile* d_tilesF = nullptr;   // image -> LP (per span)
    lpk::GTile* d_tilesP = nullptr;   // LP -> polar (per span)
    lpk::GTile* d_tilesA = nullptr;   // sinogram -> LP (per span)
    lpk::GTile* d_tilesC = nullptr;   // LP -> Cartesian (groups of 3 spans)
    int ntilesF = 0, ntilesP = 0, ntilesA = 0, ntilesC = 0;
    int boxF = 0, boxP = 0, boxA = 0, boxC = 0;
    int box_cap = 6144;   // floats of LDS per gather block (24 KiB: 6 blocks per CU)
    // box_stride rule per tile list, bit k for image->LP, LP->polar,
    // sinogram->LP, LP->Cartesian: the bank-spread stride measured faster for
    // the sinogram->LP gather only (1.00 -> 0.92 ms per 32 slices at N = 2048)
    int box_rule = 4;
    int stop_after = 0;   // stage-wise testing: return after this stage (0 = never)

    float* d_Rc = nullptr;
    float2* d_cmapA = nullptr;
    int* d_xrow = nullptr;      // [3][Nrho+1] offsets of each (span, row)'s extra samples in d_xent
    float4* d_xent = nullptr;   // (theta, x, y, row) of the extra samples
    int nxent = 0;
    float2* d_cmapC = nullptr;
    float2* d_fzA = nullptr;
    float* d_H2 = nullptr;

    // stage timing: one set of kNStage+1 events per profiled call
    static constexpr int kNStage = 7;
    bool prof = false;
    std::vector<std::vector<hipEvent_t>> prof_sets[2];
    size_t prof_used[2] = {0, 0};

    hipEvent_t* prof_begin(int which) {
        if (!prof) return nullptr;
        auto& sets = prof_sets[which];
        if (prof_used[which] == sets.size()) {
            std::vector<hipEvent_t> ev(kNStage + 1);
            for (auto& e : ev)
                if (hipEventCreate(&e) != hipSuccess) return nullptr;
            sets.push_back(ev);
        }
        return sets[prof_used[which]++].data();
    }
    void prof_clear() {
        for (int w = 0; w < 2; ++w) {
            for (auto& set : prof_sets[w])
                for (auto& e : set) (void)hipEventDestroy(e);
            prof_sets[w].clear();
            prof_used[w] = 0;
        }
    }

    int fail(int code, const std::string& msg) {
        err = msg;
        return code;
    }

    template <class T>
    int alloc(T** p, size_t count) {
        void* q = nullptr;
        size_t nb = std::max<size_t>(count, 1) * sizeof(T);
        hipError_t e = hipMalloc(&q, nb);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail(LP_ERR_ALLOC, std::string("hipMalloc of ") + std::to_string(nb) + " bytes failed: " +
                                          hipGetErrorString(e));
        }
        allocs.push_back(q);
        bytes += nb;
        *p = (T*)q;
        return LP_OK;
    }
    template <class T>
    int upload(T** p, const std::vector<T>& v) {
        int rc = alloc(p, v.size());
        if (rc) return rc;
        if (v.empty()) return LP_OK;
        hipError_t e = hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
        if (e != hipSuccess) return fail(LP_ERR_DEVICE, std::string("hipMemcpy: ") + hipGetErrorString(e));
        return LP_OK;
    }
    int twiddles(int M) {
        if (tw.count(M)) return LP_OK;
        std::vector<float2> t(M);
        for (int j = 0; j < M; ++j) {
            double a = -2.0 * M_PI * (double)j / (double)M;
            t[j] = make_float2((float)std::cos(a), (float)std::sin(a));
        }
        float2* d = nullptr;
        int rc = upload(&d, t);
        if (rc) return rc;
        tw[M] = d;
        return LP_OK;
    }
    ~lp_handle() {
        DeviceGuard g(device);
        prof_clear();
        if (rtc_mod) (void)hipModuleUnload(rtc_mod);
        for (void* p : allocs) (void)hipFree(p);
    }
};

namespace {

int check_launch(lp_handle* h, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return h->fail(LP_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
    return LP_OK;
}

inline bool valid2(float2 v) { return std::isfinite(v.x) && std::isfinite(v.y); }

// base rows per band of the fused inverse row pass + polar gather (rows per
// block 16, minus the 3 extra rows of the cubic taps)
constexpr int kFuseBase = 13;

// Source box of one output tile: every cubic tap (floor(x)-1 .. floor(x)+2)
// of every valid point, in unwrapped source coordinates.
template <class Coord>
bool tile_box(Coord coord, int span, int r0, int c0, int nr, int nc, int& x0, int& y0, int& bw, int& bh) {
    long long xmin = LLONG_MAX, xmax = LLONG_MIN, ymin = LLONG_MAX, ymax = LLONG_MIN;
    for (int r = 0; r < nr; ++r)
        for (int c = 0; c < nc; ++c) {
            float2 v = coord(span, r0 + r, c0 + c);
            if (!valid2(v)) continue;
            long long ix = (long long)std::floor(v.x), iy = (long long)std::floor(v.y);
            xmin = std::min(xmin, ix - 1); xmax = std::max(xmax, ix + 2);
            ymin = std::min(ymin, iy - 1); ymax = std::max(ymax, iy + 2);
        }
    if (xmin > xmax) { x0 = y0 = bw = bh = 0; return false; }
    if (xmax - xmin > (1 << 20) || ymax - ymin > (1 << 20)) throw std::runtime_error("coordinate table out of range");
    x0 = (int)xmin; y0 = (int)ymin; bw = (int)(xmax - xmin + 1); bh = (int)(ymax - ymin + 1);
    return true;
}

// LDS row stride of a gather box of width bw.  Rule 0: bw | 1.  Rule 1 (bank
// spread): the least s >= bw whose residue makes every nonzero offset
// (dx, dy), dy >= 0, with dx + dy s = 0 (mod 32) -- two box slots on one LDS
// bank in the 32-bank view, a fortiori in the 64-bank one -- at least ~5 slots
// long (|.|^2 >= 26), and >= 34 for dx + dy s = 0 (mod 64): the lanes of a
// wave read taps of nearby points, so short same-bank offsets are the
// conflicts that occur.  Odd strides include s = 1 (mod 32), whose offset
// (-1, 1) puts diagonal neighbours on one bank.
int box_stride(int bw, int rule) {
    if (rule == 0 || bw <= 0) return bw | 1;
    static const std::vector<char> good = [] {
        auto shortest = [](int s, int B) {
            int best = INT_MAX;
            for (int dy = 0; dy <= 16; ++dy)
                for (int dx = -40; dx <= 40; ++dx) {
                    if (dy == 0 && dx <= 0) continue;
                    if (((dx + dy * s) % B + B) % B == 0) best = std::min(best, dx * dx + dy * dy);
                }
            return best;
        };
        std::vector<char> g(64);
        for (int s = 0; s < 64; ++s) g[s] = shortest(s, 32) >= 26 && shortest(s, 64) >= 34;
        return g;
    }();
    int s = bw;
    while (!good[s & 63]) ++s;
    return s;
}

// Tiles of the output rectangle rows [rb, re) x cols [0, nc_all), split until
// every span's box fits `cap` floats (column counts stay powers of two: the
// rows are halved, or the columns).  `spans` tiles are emitted per output
// tile (one per span, consecutive).  keep(r0, c0, nr, nc) may drop a tile.
template <class Coord, class Keep>
void add_tiles(std::vector<lpk::GTile>& out, Coord coord, const std::vector<int>& spans, int r0, int c0, int nr,
               int nc, int cap, int rule, Keep keep, int& maxbox) {
    if (!keep(r0, c0, nr, nc)) return;
    if (nc <= 0 || (nc & (nc - 1)) || nc > lpk::kGatherThreads || (long long)nr * nc > lpk::kGatherThreads * lpk::kGatherPPT)
        throw std::runtime_error("gather tile shape out of range");
    std::vector<lpk::GTile> t(spans.size());
    bool fits = true;
    for (size_t i = 0; i < spans.size(); ++i) {
        lpk::GTile& g = t[i];
        g.r0 = r0; g.c0 = c0; g.nr = nr; g.nc = nc; g.span = spans[i]; g.lnc = __builtin_ctz((unsigned)nc);
        g.cr = g.cc = 0;
        g.pbase = g.cbase = 0;
        tile_box(coord, spans[i], r0, c0, nr, nc, g.x0, g.y0, g.bw, g.bh);
        g.bs = box_stride(g.bw, rule);
        if ((long long)g.bs * (g.bh | 1) > cap) fits = false;
    }
    if (!fits) {
        if (nr == 1 && nc == 1) throw std::runtime_error("a single sample exceeds the gather box budget");
        if (nr >= nc) {
            int h = nr / 2;
            add_tiles(out, coord, spans, r0, c0, h, nc, cap, rule, keep, maxbox);
            add_tiles(out, coord, spans, r0 + h, c0, nr - h, nc, cap, rule, keep, maxbox);
        } else {
            int h = nc / 2;
            add_tiles(out, coord, spans, r0, c0, nr, h, cap, rule, keep, maxbox);
            add_tiles(out, coord, spans, r0, c0 + h, nr, nc - h, cap, rule, keep, maxbox);
        }
        return;
    }
    for (auto& g : t) {
        maxbox = std::max(maxbox, g.bs * (g.bh | 1));
        out.push_back(g);
    }
}

// Colour classes of a tile for the transposed (scatter) use: the smallest
// (cr, cc) found such that points (r, c) and (r', c') with r = r' (mod cr)
// and c = c' (mod cc) never share a cubic tap (their floor coordinates differ
// by 4 or more in x or in y).  Nearest clashes along each axis set the
// periods; same-colour neighbours up to two periods away are then checked
// and a period is widened until none clash.  More than 256 classes (or none
// found): cr = 0, the scatter accumulates with atomics.
template <class Coord>
void colour_tile(lpk::GTile& t, Coord coord) {
    const int nr = t.nr, nc = t.nc;
    std::vector<int> fx((size_t)nr * nc), fy((size_t)nr * nc);
    for (int r = 0; r < nr; ++r)
        for (int c = 0; c < nc; ++c) {
            const float2 v = coord(t.span, t.r0 + r, t.c0 + c);
            const bool ok = valid2(v);
            fx[(size_t)r * nc + c] = ok ? (int)std::floor(v.x) : INT_MIN;
            fy[(size_t)r * nc + c] = ok ? (int)std::floor(v.y) : 0;
        }
    auto clash = [&](int r1, int c1, int r2, int c2) {
        if (r2 < 0 || r2 >= nr || c2 < 0 || c2 >= nc) return false;
        const size_t i = (size_t)r1 * nc + c1, j = (size_t)r2 * nc + c2;
        if (fx[i] == INT_MIN || fx[j] == INT_MIN) return false;
        return std::abs(fx[i] - fx[j]) <= 3 && std::abs(fy[i] - fy[j]) <= 3;
    };
    int cr = 1, cc = 1;
    for (int r = 0; r < nr; ++r)
        for (int c = 0; c < nc; ++c) {
            for (int d = 1; d < std::min(nr, 64); ++d)
                if (clash(r, c, r + d, c)) cr = std::max(cr, d + 1);
            for (int d = 1; d < std::min(nc, 64); ++d)
                if (clash(r, c, r, c + d)) cc = std::max(cc, d + 1);
        }
    for (int iter = 0; iter < 64 && cr * cc <= 256; ++iter) {
        bool bad = false;
        for (int r = 0; r < nr && !bad; ++r)
            for (int c = 0; c < nc && !bad; ++c)
                for (int a = -2; a <= 2 && !bad; ++a)
                    for (int b = -2; b <= 2 && !bad; ++b)
                        if ((a || b) && clash(r, c, r + a * cr, c + b * cc)) {
                            bad = true;
                            if (cr <= cc) ++cr; else ++cc;
                        }
        if (!bad) {
            t.cr = std::min(cr, nr);
            t.cc = std::min(cc, nc);
            return;
        }
    }
    t.cr = t.cc = 0;
}

// Colours every tile of a list (threads over tiles) and lays out the
// colour-sorted point lists: perm (r * nc + c of the points with a sample,
// stable within a class) and the class bounds cls, both concatenated over the
// tiles (GTile.pbase / cbase); uploaded to *d_perm / *d_cls.
template <class Coord>
int colour_list(lp_handle* h, std::vector<lpk::GTile>& tiles, Coord coord, unsigned short** d_perm, int** d_cls) {
    const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::vector<unsigned short>> perms(tiles.size());
    std::vector<std::vector<int>> clss(tiles.size());
    std::vector<std::thread> pool;
    for (unsigned w = 0; w < nth; ++w)
        pool.emplace_back([&, w] {
            for (size_t i = w; i < tiles.size(); i += nth) {
                lpk::GTile& t = tiles[i];
                colour_tile(t, coord);
                if (t.cr <= 0) continue;
                const int ncol = t.cr * t.cc;
                std::vector<int> cnt(ncol + 1, 0);
                std::vector<int> col((size_t)t.nr * t.nc, -1);
                for (int r = 0; r < t.nr; ++r)
                    for (int c = 0; c < t.nc; ++c)
                        if (valid2(coord(t.span, t.r0 + r, t.c0 + c))) {
                            const int k = (r % t.cr) * t.cc + c % t.cc;
                            col[(size_t)r * t.nc + c] = k;
                            cnt[k + 1]++;
                        }
                for (int k = 0; k < ncol; ++k) cnt[k + 1] += cnt[k];
                std::vector<unsigned short> pm(cnt[ncol]);
                std::vector<int> at(cnt.begin(), cnt.end() - 1);
                for (size_t p = 0; p < col.size(); ++p)
                    if (col[p] >= 0) pm[at[col[p]]++] = (unsigned short)p;
                perms[i] = std::move(pm);
                clss[i] = std::move(cnt);
            }
        });
    for (auto& th : pool) th.join();
    std::vector<unsigned short> perm;
    std::vector<int> cls;
    for (size_t i = 0; i < tiles.size(); ++i) {
        tiles[i].pbase = (int)perm.size();
        tiles[i].cbase = (int)cls.size();
        perm.insert(perm.end(), perms[i].begin(), perms[i].end());
        cls.insert(cls.end(), clss[i].begin(), clss[i].end());
    }
    int rc;
    if ((rc = h->upload(d_perm, perm)) || (rc = h->upload(d_cls, cls))) return rc;
    return LP_OK;
}

template <class Coord, class Keep>
std::vector<lpk::GTile> tile_grid(Coord coord, const std::vector<int>& spans, int rb, int re, int ncols, int TR, int TC,
                                  int cap, int rule, Keep keep, int& maxbox) {
    std::vector<lpk::GTile> out;
    for (int r = rb; r < re; r += TR)
        for (int c = 0, w; c < ncols; c += w) {
            // TC is a power of two; a ragged last column block is cut into
            // power-of-two widths
            w = std::min(TC, ncols - c);
            while (w & (w - 1)) w &= w - 1;
            add_tiles(out, coord, spans, r, c, std::min(TR, re - r), w, cap, rule, keep, maxbox);
        }
    return out;
}

// [lo, hi) theta columns of each (span, row) of the log-polar grids that the
// source boxes of a gather's tiles cover; a row whose boxes wrap around the
// periodic theta axis gets the whole row
std::vector<int2> box_row_ranges(const std::vector<lpk::GTile>& tiles, int Nr, int Nt) {
    std::vector<int2> r(3 * (size_t)Nr, make_int2(INT_MAX, INT_MIN));
    std::vector<char> full(3 * (size_t)Nr, 0);
    for (const auto& t : tiles) {
        if (t.bw <= 0) continue;
        bool wraps = t.x0 < 0 || t.x0 + t.bw > Nt;
        int ny = std::min(t.bh, Nr);
        for (int yy = 0; yy < ny; ++yy) {
            int row = ((t.y0 + yy) % Nr + Nr) % Nr;
            size_t k = (size_t)t.span * Nr + row;
            if (wraps) full[k] = 1;
            r[k].x = std::min(r[k].x, t.x0);
            r[k].y = std::max(r[k].y, t.x0 + t.bw);
        }
        if (t.bh > Nr) for (int row = 0; row < Nr; ++row) full[(size_t)t.span * Nr + row] = 1;
    }
    for (size_t k = 0; k < r.size(); ++k) {
        if (full[k]) r[k] = make_int2(0, Nt);
        else if (r[k].x > r[k].y) r[k] = make_int2(0, 0);
    }
    return r;
}

// [lo, hi) theta bounds of the non-empty points of each (span, row) of a dense map
std::vector<int2> row_ranges(const std::vector<float2>& cmap, int Nr, int Nt) {
    std::vector<int2> r(3 * (size_t)Nr);
    for (size_t k = 0; k < 3; ++k)
        for (int i = 0; i < Nr; ++i) {
            const float2* row = cmap.data() + (k * Nr + i) * (size_t)Nt;
            int lo = Nt, hi = 0;
            for (int j = 0; j < Nt; ++j)
                if (row[j].x == row[j].x) { lo = std::min(lo, j); hi = j + 1; }
            r[k * Nr + i] = make_int2(lo < hi ? lo : 0, lo < hi ? hi : 0);
        }
    return r;
}

// Per span, the rows [lo, hi) outside which a gathered grid is empty (from
// the per-row sampled ranges), widened to multiples of 128 (the column pass
// skips whole 128-row chunks).  Empty when every span needs all rows.
std::vector<int2> data_rows(const std::vector<int2>& rng, int Nr) {
    std::vector<int2> out(3);
    bool all = true;
    for (int k = 0; k < 3; ++k) {
        int lo = Nr, hi = 0;
        for (int r = 0; r < Nr; ++r)
            if (rng[(size_t)k * Nr + r].x < rng[(size_t)k * Nr + r].y) { lo = std::min(lo, r); hi = r + 1; }
        if (lo >= hi) lo = hi = 0;
        lo = lo / 128 * 128;
        hi = std::min(Nr, (hi + 127) / 128 * 128);
        out[k] = make_int2(lo, hi);
        if (lo != 0 || hi != Nr) all = false;
    }
    if (all) out.clear();
    return out;
}

// Rows each output gather reads (non-empty read ranges); empty when all are read.
std::vector<unsigned char> read_rows(const std::vector<int2>& wrng) {
    std::vector<unsigned char> out(wrng.size());
    bool all = true;
    for (size_t i = 0; i < wrng.size(); ++i) {
        out[i] = wrng[i].x < wrng[i].y;
        if (!out[i]) all = false;
    }
    if (all) out.clear();
    return out;
}

// kernel spectrum [Nrho][H] interleaved (wire format) -> transposed [H][Nrho]
// with the C = 1/(Ntheta Nrho) of the reference's mul kernel folded in.
std::vector<float2> transpose_fz(const float* src, int Nr, int Nt) {
    int H = Nt / 2 + 1;
    double sc = 1.0 / ((double)Nt * (double)Nr);
    std::vector<float2> out((size_t)H * Nr);
    for (int i = 0; i < Nr; ++i)
        for (int c = 0; c < H; ++c) {
            double re = src[2 * ((size_t)i * H + c)], im = src[2 * ((size_t)i * H + c) + 1];
            out[(size_t)c * Nr + i] = make_float2((float)(re * sc), (float)(im * sc));
        }
    return out;
}

// LDS line stride of the row passes (bank spreading is done by lidx())
int row_ldm(const lp_handle* h, int R) {
    (void)R;
    return h->h;
}

int pick_rows(lp_handle* h) {
    // log-polar rows per row-pass block: as many as the LDS and the FFT
    // register budget allow, up to the tuned value
    int R = h->R;
    while (R > 1) {
        size_t lds = lpk::rows_lds_bytes(row_ldm(h, R), R);
        long long pts = (long long)R * h->h;
        if (lds <= h->lds_limit && pts <= 32LL * 1024 && h->Nr % R == 0) break;
        R /= 2;
    }
    return R;
}

// Scratch of lp_fwd_noalloc: the device buffers of one forward call, carved
// from one statically allocated array per GPU so that the call allocates
// nothing.  Each GPU the library runs on holds its own copy until the process
// ends.  In host code `scratch` names a host-side stand-in: its device address
// comes from hipGetSymbolAddress, once per device.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on `device` (which must be current).
int scratch_address(int device, void** out) {
    if (device < 0 || device >= kMaxDevices) return LP_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            (void)hipGetLastError();
            return LP_ERR_DEVICE;
        }
        g_scratch_dev[device] = p;
    }
    *out = g_scratch_dev[device];
    return LP_OK;
}

inline size_t round256(size_t b) { return (b + 255) & ~size_t(255); }

// LPREC_AMD_VERBOSE=1: one line per tile list on stderr at init (tiles,
// coloured tiles and their mean / largest class count, staged floats per slice).
void report_tiles(const char* name, const std::vector<lpk::GTile>& tiles, long long staged) {
    const char* v = std::getenv("LPREC_AMD_VERBOSE");
    if (!v || v[0] != '1') return;
    size_t coloured = 0;
    long long sum = 0, mx = 0;
    for (const auto& t : tiles)
        if (t.cr > 0) {
            ++coloured;
            sum += (long long)t.cr * t.cc;
            mx = std::max(mx, (long long)t.cr * t.cc);
        }
    std::fprintf(stderr, "lprec_amd: %s: %zu tiles, %zu coloured (mean %.1f, max %lld classes), %lld staged floats/slice\n",
                 name, tiles.size(), coloured, coloured ? (double)sum / coloured : 0.0, mx, staged);
}

// Staging tables of a tile list (launch_scatter with Scatter.stage, then
// launch_box_sum): box i = tiles[i] (tile * group + span slot), its offset in
// a slice's staging area, and the destination tiles (32 x 32 points of a
// W x H plane; plane = the box's span when per_span) that it covers with
// periodic placement, listed per destination tile in increasing box order.
int stage_tables(lp_handle* h, const std::vector<lpk::GTile>& tiles, int W, int H, bool per_span,
                 lp_handle::Stage& st) {
    const int tx = (W + 31) / 32, ty = (H + 31) / 32, np = per_span ? 3 : 1;
    std::vector<int> off(tiles.size());
    std::vector<int4> rect(tiles.size());
    std::vector<std::vector<int>> lists((size_t)np * tx * ty);
    long long total = 0;
    for (size_t i = 0; i < tiles.size(); ++i) {
        const lpk::GTile& t = tiles[i];
        off[i] = (int)total;
        rect[i] = make_int4(t.x0, t.y0, std::max(t.bw, 0), std::max(t.bh, 0));
        if (t.bw <= 0 || t.bh <= 0) continue;
        // a box wider than the plane, or an origin beyond one period: keep the atomic flush
        if (t.bw > W || t.bh > H || t.x0 <= -W || t.x0 >= W || t.y0 <= -H || t.y0 >= H) return LP_OK;
        total += (long long)t.bw * t.bh;
        if (total >= (1LL << 31)) return LP_OK;
        const int plane = per_span ? t.span : 0;
        // destination tile columns / rows the wrapped box touches
        auto cover = [](int o, int n, int L, int tile, std::vector<int>& out) {
            std::vector<char> hit((L + tile - 1) / tile, 0);
            for (int d = 0; d < n; ++d) hit[(((o + d) % L) + L) % L / tile] = 1;
            out.clear();
            for (int k = 0; k < (int)hit.size(); ++k)
                if (hit[k]) out.push_back(k);
        };
        std::vector<int> cx, cy;
        cover(t.x0, t.bw, W, 32, cx);
        cover(t.y0, t.bh, H, 32, cy);
        for (int yy : cy)
            for (int xx : cx) lists[((size_t)plane * ty + yy) * tx + xx].push_back((int)i);
    }
    std::vector<int> tbeg(lists.size() + 1, 0), tbox;
    for (size_t k = 0; k < lists.size(); ++k) {
        tbox.insert(tbox.end(), lists[k].begin(), lists[k].end());
        tbeg[k + 1] = (int)tbox.size();
    }
    int rc;
    if ((rc = h->upload(&st.off, off)) || (rc = h->upload(&st.rect, rect)) || (rc = h->upload(&st.tbeg, tbeg)) ||
        (rc = h->upload(&st.tbox, tbox)))
        return rc;
    st.total = total;
    st.tiles_x = tx;
    st.tiles_y = ty;
    st.nplanes = np;
    st.ok = true;
    return LP_OK;
}

// Tables of the separable polar transpose (launch_polar_transpose): the
// angle range of each span, each kept angle's theta coordinate, the first
// detector cell per angle and unwrapped rho row, the first angle per span and
// theta column.  Leaves pt_ok false when the polar map is not separable in
// the way the kernels need (a theta coordinate that varies along the
// detector, rho not monotone, an angle whose taps wrap around rho more than
// once, taps that wrap around the theta axis).
int polar_transpose_tables(lp_handle* h, const std::vector<float2>& p2lp, const std::vector<unsigned char>& spanrow) {
    const int N = h->N, Nr = h->Nr, Nt = h->Nt, P0 = h->P0, N0 = h->N0, os = h->os;
    const int lb = N / 2 - h->cor;
    auto c_of = [&](int a, int x) { return p2lp[(size_t)a * os * N + x + lb]; };
    std::vector<int> abeg(6, 0);
    for (int k = 0; k < 3; ++k) {
        int ab = -1, ae = -1;
        for (int a = 0; a < P0; ++a)
            if (spanrow[(size_t)a * os] == k) { if (ab < 0) ab = a; ae = a + 1; }
        if (ab < 0) ab = ae = 0;
        abeg[k] = ab;
        abeg[3 + k] = ae;
    }
    std::vector<float> cxa(P0);
    long long qmin = LLONG_MAX, qmax = LLONG_MIN;
    for (int a = 0; a < P0; ++a) {
        cxa[a] = c_of(a, 0).x;
        long long prev = LLONG_MIN;
        for (int x = 0; x < N0; ++x) {
            const float2 c = c_of(a, x);
            if (c.x != cxa[a]) return LP_OK;
            const long long iy = (long long)std::floor(c.y);
            if (iy < prev) return LP_OK;
            prev = iy;
        }
        // an angle's taps may wrap around rho once (the span edges reach
        // across the whole grid): at most two unwrapped rows per grid row
        const long long i0 = (long long)std::floor(c_of(a, 0).y), i1 = prev;
        if (i1 - i0 + 4 >= 2LL * Nr) return LP_OK;
        qmin = std::min(qmin, i0);
        qmax = std::max(qmax, i1);
    }
    const long long qlo = qmin - 3, Q = qmax + 3 - qlo;
    if (Q > 4LL * Nr) return LP_OK;
    std::vector<int> xs((size_t)P0 * (Q + 1));
    for (int a = 0; a < P0; ++a) {
        int x = 0;
        for (long long i = 0; i <= Q; ++i) {
            while (x < N0 && (long long)std::floor(c_of(a, x).y) < qlo + i) ++x;
            xs[(size_t)a * (Q + 1) + i] = x;
        }
    }
    long long imin = LLONG_MAX, imax = LLONG_MIN;
    for (int k = 0; k < 3; ++k)
        for (int a = abeg[k]; a < abeg[3 + k]; ++a) {
            const long long ix = (long long)std::floor(cxa[a]);
            if (a > abeg[k] && ix < (long long)std::floor(cxa[a - 1])) return LP_OK;
            imin = std::min(imin, ix);
            imax = std::max(imax, ix);
        }
    if (imin > imax) return LP_OK;
    for (int k = 0; k < 3; ++k)   // the intermediate [3][Nr][angles] must fit the half-spectrum buffer
        if (abeg[3 + k] - abeg[k] > Nt) return LP_OK;
    // the cubic tap extent also for linear sampling (zeros in its outer
    // columns): it is the extent of the gather boxes, the columns the row pass
    // then reads, so every one of them must be written
    const long long jlo = imin - 1, jhi = imax + 3;
    if (jlo < 0 || jhi > Nt) return LP_OK;
    const int jn = (int)(jhi - jlo);
    std::vector<int> aj(3 * (size_t)(jn + 5));
    int nA = 1;
    for (int k = 0; k < 3; ++k) {
        const int na = abeg[3 + k] - abeg[k];
        nA = std::max(nA, na);
        int al = 0;
        for (int i = 0; i <= jn + 4; ++i) {
            while (al < na && (long long)std::floor(cxa[abeg[k] + al]) < jlo - 2 + i) ++al;
            aj[(size_t)k * (jn + 5) + i] = al;
        }
    }
    int rc;
    if ((rc = h->upload(&h->d_pt_abeg, abeg)) || (rc = h->upload(&h->d_pt_cxa, cxa)) ||
        (rc = h->upload(&h->d_pt_xs, xs)) || (rc = h->upload(&h->d_pt_aj, aj)))
        return rc;
    h->pt_nA = nA;
    h->pt_qlo = (int)qlo;
    h->pt_Q = (int)Q;
    h->pt_jlo = (int)jlo;
    h->pt_jn = jn;
    h->pt_ok = true;
    return LP_OK;
}

// ---- row / column passes compiled at run time ------------------------------
// For lengths with odd factors (a centre offset makes N = 2064 = 2^4 3 43 at
// N0 = 2048) the generic mixed-radix kernels spend most of their time on run-
// time index arithmetic.  rtc_build instantiates rtc_kernels.hpp (the same
// passes with every size a template argument) for this handle's lengths and
// radix sequences with hipRTC, once per process and set of lengths, and loads
// the code object on the handle's device.  Failure is not an error: the
// generic kernels stay in use (LPREC_AMD_VERBOSE=1 prints why).

// directory of the device sources next to the shared library
std::string csrc_dir() {
    Dl_info info{};
    if (!dladdr((void*)&lp_create, &info) || !info.dli_fname) return "";
    std::string p(info.dli_fname);
    const size_t k = p.find_last_of('/');
    return (k == std::string::npos ? std::string(".") : p.substr(0, k)) + "/csrc";
}

std::string radix_list(const FftPlan& p) {
    std::string r;
    for (int i = 0; i < p.nst; ++i) r += ", " + std::to_string(p.radix[i]);
    return r;
}

struct RtcCode {
    std::vector<char> code;
    std::vector<std::string> lowered;
};
std::mutex g_rtc_mu;
std::map<std::string, std::shared_ptr<RtcCode>> g_rtc_cache;

std::shared_ptr<RtcCode> rtc_compile(const std::string& arch, int pp, const std::vector<std::string>& names,
                                     std::string& err) {
    std::string key = arch + ";" + std::to_string(pp);
    for (const auto& n : names) key += ";" + n;
    std::lock_guard<std::mutex> lock(g_rtc_mu);
    auto it = g_rtc_cache.find(key);
    if (it != g_rtc_cache.end()) return it->second;
    const std::string dir = csrc_dir();
    if (dir.empty()) { err = "cannot locate the device sources"; return nullptr; }
    const std::string src = "#include \"rtc_kernels.hpp\"\n";
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "lprec_rtc.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
        err = "hiprtcCreateProgram failed";
        return nullptr;
    }
    for (const auto& n : names) hiprtcAddNameExpression(prog, n.c_str());
    const std::string a = "--offload-arch=" + arch, inc = "-I" + dir, dpp = "-DLPCT_ODD_PP=" + std::to_string(pp);
    // packed FP32 instructions off, as for the library's own kernels
    const char* opts[] = {a.c_str(), "-O3", "-std=c++17", inc.c_str(), dpp.c_str(),
                          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    if (hiprtcCompileProgram(prog, 9, opts) != HIPRTC_SUCCESS) {
        size_t n = 0;
        hiprtcGetProgramLogSize(prog, &n);
        std::string log(n, '\0');
        if (n) hiprtcGetProgramLog(prog, &log[0]);
        err = "hipRTC compile failed: " + log.substr(0, 2000);
        hiprtcDestroyProgram(&prog);
        return nullptr;
    }
    auto out = std::make_shared<RtcCode>();
    for (const auto& n : names) {
        const char* low = nullptr;
        if (hiprtcGetLoweredName(prog, n.c_str(), &low) != HIPRTC_SUCCESS || !low) {
            err = "no lowered name for " + n;
            hiprtcDestroyProgram(&prog);
            return nullptr;
        }
        out->lowered.emplace_back(low);
    }
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    out->code.resize(cs);
    if (!cs || hiprtcGetCode(prog, out->code.data()) != HIPRTC_SUCCESS) {
        err = "hiprtcGetCode failed";
        hiprtcDestroyProgram(&prog);
        return nullptr;
    }
    hiprtcDestroyProgram(&prog);
    g_rtc_cache[key] = out;
    return out;
}

bool pow2_in(int n, int lo, int hi) { return n >= lo && n <= hi && (n & (n - 1)) == 0; }

// Output pairs per odd-radix work item (LPCT_ODD_PP of fft_ct.hpp, one value
// per compiled module) from 4, 6 and 8: the fewest cycles over the module's
// passes, counting for each radix R > 5 the rounds of the block's threads
// over its work items (lines x M/R butterflies x ceil(hR/PP) groups) times
// the pairs per item (hR = (R-1)/2 terms each).  E.g. the 4128-point column
// at 512 threads: PP 4 -> 576 items, two rounds of 4; PP 6 -> 384, one of 6.
int odd_pairs(const std::vector<std::array<int, 4>>& passes, const std::vector<const FftPlan*>& plans) {
    int best = 4;
    long long best_cost = LLONG_MAX;
    for (int pp : {4, 6, 8}) {
        long long cost = 0;
        for (size_t k = 0; k < passes.size(); ++k) {
            const int M = passes[k][0], NL = passes[k][1], NT = passes[k][2], weight = passes[k][3];
            const FftPlan& pl = *plans[k];
            for (int i = 0; i < pl.nst; ++i) {
                const int R = pl.radix[i];
                if (R <= 5 || !(R & 1)) continue;
                const int hR = (R - 1) / 2;
                const int P = hR <= 2 ? 2 : (hR <= pp ? ((hR + 1) & ~1) : pp);
                const long long items = (long long)NL * (M / R) * ((hR + P - 1) / P);
                cost += weight * ((items + NT - 1) / NT) * P * hR;
            }
        }
        if (cost < best_cost) { best_cost = cost; best = pp; }
    }
    return best;
}

// threads of a mixed-radix pass over `points` values (as threads_for)
int rtc_threads(long long points) {
    int nt = 256;
    while (points > 16LL * nt && nt < 1024) nt *= 2;
    return nt;
}

void rtc_build(lp_handle* h) {
    if (h->rtc_tried || !h->rtc_mode) return;
    h->rtc_tried = true;
    if (h->rtc_mod) {
        (void)hipModuleUnload(h->rtc_mod);
        h->rtc_mod = nullptr;
        h->rtc_fn[0] = h->rtc_fn[1] = h->rtc_fn[2] = h->rtc_fn[3] = nullptr;
    }
    const bool rows = !pow2_in(h->h, 256, 4096), cols = !pow2_in(h->Nr, 256, 4096);
    // the filtered pair pass has no register kernel at any length
    const bool pair = h->has_filter;
    if (!rows && !cols && !pair) return;
    const char* v = std::getenv("LPREC_AMD_VERBOSE");
    const bool verbose = v && v[0] == '1';
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, h->device) != hipSuccess) return;
    std::string arch(prop.gcnArchName);
    arch = arch.substr(0, arch.find(':'));
    // rows per block of the compiled row passes: at most 8 (the generic
    // kernels' limit) and few enough that their two LDS buffers fit
    int R = std::min(h->R, 8);
    while (R > 1 && 2 * (size_t)lds_slots(R * row_ldm(h, R)) * sizeof(float2) > h->lds_limit) R /= 2;
    const int ldm = row_ldm(h, R);
    const int rnt = rtc_threads((long long)R * h->h);
    const int cnt = h->rtc_cols_req ? h->rtc_cols_req : rtc_threads(h->Nr);
    const std::string hr = radix_list(h->plan_h), mr = radix_list(h->plan_M);
    const std::string wpe = (h->rtc_cols_var & 4) ? "4" : "2";
    const int pnt = rtc_threads(2LL * h->N);
    // only the passes this handle runs (slot: 0 rows fwd, 1 rows inv, 2 cols,
    // 3 filtered sinogram pair)
    std::vector<std::string> names;
    std::vector<int> slot;
    if (rows) {
        const std::string args = std::to_string(h->h) + ", " + std::to_string(R) + ", " + std::to_string(ldm) + ", " +
                                 std::to_string(rnt) + ", " + wpe + ", 1" + hr + ">";
        names.push_back("lpct::k_rows_fwd_ct<" + args);
        slot.push_back(0);
        names.push_back("lpct::k_rows_inv_ct<" + args);
        slot.push_back(1);
    }
    if (cols) {
        names.push_back("lpct::k_cols_ct<" + std::to_string(h->Nr) + ", " + std::to_string(cnt) + ", " + wpe + ", " +
                        std::to_string(h->rtc_cols_var & 3) + mr + ">");
        slot.push_back(2);
    }
    if (pair) {
        names.push_back("lpct::k_sino_rows_pair_ct<" + std::to_string(h->N) + ", " + std::to_string(pnt) + ", " + wpe +
                        radix_list(h->plan_2N) + ">");
        slot.push_back(3);
    }
    // the passes of the module (length, lines per block, threads, launches
    // per fwd or adj call) for the choice of output pairs per work item
    std::vector<std::array<int, 4>> passes;
    std::vector<const FftPlan*> plans;
    if (rows) {
        passes.push_back({h->h, R, rnt, 2});
        plans.push_back(&h->plan_h);
    }
    if (cols) {
        passes.push_back({h->Nr, 1, cnt, 1});
        plans.push_back(&h->plan_M);
    }
    if (pair) {
        passes.push_back({2 * h->N, 1, pnt, 1});
        plans.push_back(&h->plan_2N);
    }
    std::string err;
    const int pp = (h->rtc_cols_var & 8) ? odd_pairs(passes, plans) : 8;
    auto code = rtc_compile(arch, pp, names, err);
    hipModule_t mod = nullptr;
    if (code && hipModuleLoadData(&mod, code->code.data()) != hipSuccess) {
        (void)hipGetLastError();
        err = "hipModuleLoadData failed";
        mod = nullptr;
    }
    void* fn[4] = {nullptr, nullptr, nullptr, nullptr};
    for (size_t i = 0; mod && i < names.size(); ++i) {
        hipFunction_t f = nullptr;
        if (hipModuleGetFunction(&f, mod, code->lowered[i].c_str()) != hipSuccess) {
            (void)hipGetLastError();
            err = "hipModuleGetFunction failed for " + names[i];
            (void)hipModuleUnload(mod);
            mod = nullptr;
            break;
        }
        fn[slot[i]] = (void*)f;
    }
    if (!mod) {
        if (verbose) std::fprintf(stderr, "lprec_amd: run-time compiled passes unavailable (%s)\n", err.c_str());
        return;
    }
    h->rtc_mod = mod;
    for (int i = 0; i < 4; ++i) h->rtc_fn[i] = fn[i];
    h->rtc_R = R;
    h->rtc_rows_nt = rnt;
    h->rtc_cols_nt = cnt;
    h->rtc_pair_nt = pnt;
    h->rtc_pp = pp;
    if (verbose)
        for (const auto& n : names)
            std::fprintf(stderr, "lprec_amd: run-time compiled pass: %s (odd-radix output pairs %d)\n", n.c_str(), pp);
}

}  // namespace

extern "C" {

int lp_create(const float* g, int device, lp_handle** out) {
    g_create_err.clear();
    if (!out || !g) { g_create_err = "null argument"; return LP_ERR_ARG; }
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        g_create_err = "no HIP device available";
        return LP_ERR_DEVICE;
    }
    if (device < 0 || device >= ndev) { g_create_err = "device index out of range"; return LP_ERR_ARG; }
    std::unique_ptr<lp_handle> h(new lp_handle());
    h->device = device;
    h->N = (int)g[0]; h->N0 = (int)g[1]; h->Nt = (int)g[2]; h->Nr = (int)g[3];
    int nspan = (int)g[4];
    h->P = (int)g[5]; h->Ns = (int)g[6]; h->cor = (int)g[7]; h->os = (int)g[8]; h->cubic = (int)g[9] != 0;
    if (nspan != 3 || h->N < 16 || h->N % 16 != 0 || h->Nt != h->N || h->Nr != 2 * h->N || h->os < 1 ||
        h->P % h->os != 0 || h->Ns < 1 || h->N0 < 1 || h->N0 > h->N) {
        g_create_err = "inconsistent global parameter array";
        return LP_ERR_ARG;
    }
    h->P0 = h->P / h->os;
    h->h = h->N / 2;
    int lb = h->N / 2 - h->cor;
    if (lb < 0 || lb + h->N0 > h->N) { g_create_err = "cor puts the detector outside the padded grid"; return LP_ERR_ARG; }
    DeviceGuard guard(device);
    if (!guard.ok) { g_create_err = "hipSetDevice failed"; return LP_ERR_DEVICE; }
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds > 0)
        h->lds_limit = (size_t)lds;
    h->R = 16;
    h->R = pick_rows(h.get());
    try {
        h->plan_h = make_plan(h->h);
        h->plan_N = make_plan(h->N);
        h->plan_M = make_plan(h->Nr);
    } catch (const std::exception& e) {
        g_create_err = e.what();
        return LP_ERR_UNSUPPORTED;
    }
    if (lpk::ept_for(h->Nr, 1024) < 0 || lpk::ept_for(h->N, 1024) < 0) {
        g_create_err = "grid too large for the LDS FFT";
        return LP_ERR_UNSUPPORTED;
    }
    int rc;
    size_t npts = (size_t)h->Nr * h->Nt;
    std::vector<float> erho(g + 10, g + 10 + npts);
    if ((rc = h->upload(&h->d_erho, erho)) || (rc = h->alloc(&h->d_G, 3 * (size_t)h->Ns * npts)) ||
        (rc = h->alloc(&h->d_S, 3 * (size_t)h->Ns * h->h * h->Nr)) || (rc = h->twiddles(h->h)) ||
        (rc = h->twiddles(h->N)) || (rc = h->twiddles(h->Nr))) {
        g_create_err = h->err;
        return rc;
    }
    {
        size_t need = (size_t)h->Ns * std::max((size_t)h->N * h->N, (size_t)h->P * h->N);
        if (need <= 3 * (size_t)h->Ns * npts) h->d_tmp = h->d_G;
        else if ((rc = h->alloc(&h->d_tmp, need))) { g_create_err = h->err; return rc; }
    }
    *out = h.release();
    return LP_OK;
}

int lp_init_fwd(lp_handle* h, const int32_t* pi, const float* pf) {
    if (!h) return LP_ERR_ARG;
    if (!pi || !pf) return h->fail(LP_ERR_ARG, "null parameter array");
    DeviceGuard guard(h->device);
    const int N = h->N, Nt = h->Nt, Nr = h->Nr, P = h->P;
    size_t npts = (size_t)Nr * Nt;
    // params.cu:57-99
    int Npids[3];
    size_t si = 0, sf = 0;
    for (int k = 0; k < 3; ++k) Npids[k] = pi[si++];
    std::vector<const int32_t*> pids(3);
    for (int k = 0; k < 3; ++k) { pids[k] = pi + si; si += Npids[k]; }
    int Nc = pi[si++];
    const int32_t* cids = pi + si;
    if (Nc < 0 || (size_t)Nc > npts) return h->fail(LP_ERR_ARG, "bad Ncids in fwd parameters");
    const float *c1[3], *c2[3], *q1[3], *q2[3];
    for (int k = 0; k < 3; ++k) { c1[k] = pf + sf; sf += Nc; }
    for (int k = 0; k < 3; ++k) { c2[k] = pf + sf; sf += Nc; }
    for (int k = 0; k < 3; ++k) { q1[k] = pf + sf; sf += Npids[k]; }
    for (int k = 0; k < 3; ++k) { q2[k] = pf + sf; sf += Npids[k]; }
    const float* fz = pf + sf;
    // dense per-span C->LP maps
    const float nanv = std::nanf("");
    std::vector<float2> cmap(3 * npts, make_float2(nanv, nanv));
    for (int k = 0; k < 3; ++k)
        for (int t = 0; t < Nc; ++t) {
            int id = cids[t];
            if (id < 0 || (size_t)id >= npts) return h->fail(LP_ERR_ARG, "cids index out of range");
            cmap[k * npts + id] = make_float2(c1[k][t], c2[k][t]);
        }
    // LP->polar coordinates per internal sinogram point, plus the span of each row
    std::vector<float2> p2lp((size_t)P * N, make_float2(nanv, nanv));
    std::vector<unsigned char> spanrow(P, 255);
    for (int k = 0; k < 3; ++k)
        for (int t = 0; t < Npids[k]; ++t) {
            int id = pids[k][t];
            if (id < 0 || id >= P * N) return h->fail(LP_ERR_ARG, "pids index out of range");
            p2lp[id] = make_float2(q1[k][t], q2[k][t]);
            int row = id / N;
            if (spanrow[row] != 255 && spanrow[row] != k) return h->fail(LP_ERR_UNSUPPORTED, "a sinogram row spans two spans");
            spanrow[row] = (unsigned char)k;
        }
    for (int a = 0; a < P; a += h->os)
        if (spanrow[a] == 255) return h->fail(LP_ERR_ARG, "pids do not cover every kept sinogram row");
    for (int a = 0; a < P; a += h->os)
        for (int x = 0; x < N; ++x) {
            float2 v = p2lp[(size_t)a * N + x];
            if (v.x != v.x) return h->fail(LP_ERR_ARG, "pids do not cover every kept sinogram point");
        }
    for (auto& s : spanrow) if (s == 255) s = 0;
    for (auto& v : cmap) if (!valid2(v)) v = make_float2(nanv, nanv);
    for (auto& v : p2lp) if (!valid2(v)) v = make_float2(nanv, nanv);
    std::vector<float2> fzT = transpose_fz(fz, Nr, Nt);
    int rc;
    std::vector<int2> rng = row_ranges(cmap, Nr, Nt);
    std::vector<lpk::GTile> tF, tP;
    try {
        // image -> LP gather: tiles of 32 rho x 64 theta per span, kept where
        // they meet the non-empty theta range of one of their rows
        for (int k = 0; k < 3; ++k) {
            auto coord = [&](int, int r, int c) { return cmap[(size_t)k * npts + (size_t)r * Nt + c]; };
            auto keep = [&](int r0, int c0, int nr, int nc) {
                for (int r = r0; r < r0 + nr; ++r) {
                    int2 q = rng[(size_t)k * Nr + r];
                    if (q.x < c0 + nc && q.y > c0) return true;
                }
                return false;
            };
            auto t = tile_grid(coord, {k}, 0, Nr, Nt, 64, 32, h->box_cap, (h->box_rule >> 0) & 1, keep, h->boxF);
            tF.insert(tF.end(), t.begin(), t.end());
        }
        // LP -> polar gather over the kept angles (user rows), per span
        const int P0 = h->P0, lb = N / 2 - h->cor, os = h->os, N0 = h->N0;
        for (int k = 0; k < 3; ++k) {
            int ab = -1, ae = -1;
            for (int a = 0; a < P0; ++a)
                if (spanrow[(size_t)a * os] == k) { if (ab < 0) ab = a; ae = a + 1; }
            if (ab < 0) continue;
            for (int a = ab; a < ae; ++a)
                if (spanrow[(size_t)a * os] != k) return h->fail(LP_ERR_UNSUPPORTED, "span rows are not contiguous");
            auto coord = [&](int, int a, int x) { return p2lp[(size_t)a * os * N + x + lb]; };
            auto keep = [](int, int, int, int) { return true; };
            auto t = tile_grid(coord, {k}, ab, ae, N0, 64, 32, h->box_cap, (h->box_rule >> 1) & 1, keep, h->boxP);
            tP.insert(tP.end(), t.begin(), t.end());
        }
    } catch (const std::exception& e) {
        return h->fail(LP_ERR_UNSUPPORTED, e.what());
    }
    h->ntilesF = (int)tF.size();
    h->ntilesP = (int)tP.size();
    std::vector<int2> wrng = box_row_ranges(tP, Nr, Nt);
    {   // colour classes of the tiles for the transposed (scatter) use
        const int lbp = N / 2 - h->cor, osp = h->os;
        auto coordF = [&](int k, int r, int c) { return cmap[(size_t)k * npts + (size_t)r * Nt + c]; };
        auto coordP = [&](int, int a, int x) { return p2lp[(size_t)a * osp * N + x + lbp]; };
        if ((rc = colour_list(h, tF, coordF, &h->d_permF, &h->d_clsF)) ||
            (rc = colour_list(h, tP, coordP, &h->d_permP, &h->d_clsP)))
            return rc;
    }
    if ((rc = h->upload(&h->d_tilesF, tF)) || (rc = h->upload(&h->d_tilesP, tP))) return rc;
    if ((rc = h->upload(&h->d_wrangeF, wrng))) return rc;
    {
        std::vector<int2> in_rows = data_rows(rng, Nr);
        std::vector<unsigned char> out_rows = read_rows(wrng);
        if (!in_rows.empty() && (rc = h->upload(&h->d_inrowsF, in_rows))) return rc;
        if (!out_rows.empty() && (rc = h->upload(&h->d_outrowsF, out_rows))) return rc;
    }
    if ((rc = h->upload(&h->d_cmapF, cmap)) || (rc = h->upload(&h->d_rangeF, rng)) || (rc = h->upload(&h->d_p2lp, p2lp)) ||
        (rc = h->upload(&h->d_spanrow, spanrow)) || (rc = h->upload(&h->d_fzF, fzT)) ||
        (rc = h->alloc(&h->d_fc, (size_t)h->Ns * N * N)))
        return rc;
    if ((rc = polar_transpose_tables(h, p2lp, spanrow))) return rc;
    if ((rc = stage_tables(h, tF, N, N, false, h->stF))) return rc;
    report_tiles("image->log-polar", tF, h->stF.total);
    report_tiles("log-polar->polar", tP, 0);
    {
        // polar samples of the kept angles by (span, band of kFuseBase base
        // rows): their table coordinates and output offsets a * N0 + x
        const int P0 = h->P0, lb = N / 2 - h->cor, os = h->os, N0 = h->N0;
        const int nband = (Nr + kFuseBase - 1) / kFuseBase;
        std::vector<int> cnt(3 * (size_t)(nband + 1), 0);
        std::vector<int> key((size_t)P0 * N0);
        for (int a = 0; a < P0; ++a)
            for (int x = 0; x < N0; ++x) {
                const float2 c = p2lp[(size_t)a * os * N + x + lb];
                const long long iy = (long long)std::floor(c.y);
                const int iyw = (int)(((iy % Nr) + Nr) % Nr);
                const int k = spanrow[(size_t)a * os];
                key[(size_t)a * N0 + x] = k * (nband + 1) + iyw / kFuseBase;
                cnt[key[(size_t)a * N0 + x] + 1]++;
            }
        for (size_t i = 1; i < cnt.size(); ++i) cnt[i] += cnt[i - 1];
        std::vector<int> beg(cnt);
        std::vector<float2> pc((size_t)P0 * N0);
        std::vector<int> po((size_t)P0 * N0);
        for (int a = 0; a < P0; ++a)
            for (int x = 0; x < N0; ++x) {
                const size_t at = (size_t)beg[key[(size_t)a * N0 + x]]++;
                pc[at] = p2lp[(size_t)a * os * N + x + lb];
                po[at] = a * N0 + x;
            }
        h->npband = nband;
        if ((rc = h->upload(&h->d_pcoord, pc)) || (rc = h->upload(&h->d_poff, po)) || (rc = h->upload(&h->d_pband, cnt)))
            return rc;
    }
    h->has_fwd = true;
    rtc_build(h);
    return LP_OK;
}

int lp_init_adj(lp_handle* h, const int32_t* pi, const float* pf) {
    if (!h) return LP_ERR_ARG;
    if (!pi || !pf) return h->fail(LP_ERR_ARG, "null parameter array");
    DeviceGuard guard(h->device);
    const int N = h->N, Nt = h->Nt, Nr = h->Nr;
    size_t npts = (size_t)Nr * Nt;
    // params.cu:101-145
    int Nc = pi[0], Nl = pi[1], Nw = pi[2];
    if (Nc < 0 || Nl < 0 || Nw < 0 || (size_t)Nc > (size_t)N * N || (size_t)Nl > npts || (size_t)Nw > npts)
        return h->fail(LP_ERR_ARG, "bad counts in adj parameters");
    const int32_t* cids = pi + 3;
    const int32_t* lpids = cids + Nc;
    const int32_t* wids = lpids + Nl;
    int flag = wids[Nw];
    size_t sf = 0;
    const float *C1[3], *C2[3], *L1[3], *L2[3], *W1[3], *W2[3];
    for (int k = 0; k < 3; ++k) { C1[k] = pf + sf; sf += Nc; }
    for (int k = 0; k < 3; ++k) { C2[k] = pf + sf; sf += Nc; }
    for (int k = 0; k < 3; ++k) { L1[k] = pf + sf; sf += Nl; }
    for (int k = 0; k < 3; ++k) { L2[k] = pf + sf; sf += Nl; }
    for (int k = 0; k < 3; ++k) { W1[k] = pf + sf; sf += Nw; }
    for (int k = 0; k < 3; ++k) { W2[k] = pf + sf; sf += Nw; }
    const float* fz = pf + sf;
    sf += 2 * (size_t)(Nt / 2 + 1) * Nr;
    const float* filt = flag == 1 ? pf + sf : nullptr;
    const float nanv = std::nanf("");
    // polar->LP maps: texture x = s index (lp2p2), y = angle index (lp2p1), lpRgpu.cu:307-311
    std::vector<float2> cmap(3 * npts, make_float2(nanv, nanv));
    std::vector<std::vector<float4>> extra(3 * (size_t)Nr);
    std::vector<unsigned char> used(npts, 0);
    for (int t = 0; t < Nl; ++t) {
        int id = lpids[t];
        if (id < 0 || (size_t)id >= npts) return h->fail(LP_ERR_ARG, "lpids index out of range");
        for (int k = 0; k < 3; ++k) {
            if (!used[id]) cmap[k * npts + id] = make_float2(L2[k][t], L1[k][t]);
            else extra[k * Nr + id / Nt].push_back(make_float4((float)(id % Nt), L2[k][t], L1[k][t], (float)(id / Nt)));
        }
        used[id] = 1;
    }
    for (int t = 0; t < Nw; ++t) {
        int id = wids[t];
        if (id < 0 || (size_t)id >= npts) return h->fail(LP_ERR_ARG, "wids index out of range");
        for (int k = 0; k < 3; ++k) {
            if (!used[id]) cmap[k * npts + id] = make_float2(W2[k][t], W1[k][t]);
            else extra[k * Nr + id / Nt].push_back(make_float4((float)(id % Nt), W2[k][t], W1[k][t], (float)(id / Nt)));
        }
        used[id] = 1;
    }
    std::vector<int> xrow;
    std::vector<float4> xent;
    bool any_extra = false;
    for (auto& e : extra) if (!e.empty()) any_extra = true;
    if (any_extra) {
        xrow.resize(3 * (size_t)(Nr + 1));
        for (int k = 0; k < 3; ++k) {
            xrow[k * (Nr + 1)] = (int)xent.size();
            for (int r = 0; r < Nr; ++r) {
                auto& e = extra[k * Nr + r];
                xent.insert(xent.end(), e.begin(), e.end());
                xrow[k * (Nr + 1) + r + 1] = (int)xent.size();
            }
        }
    }
    // LP->Cartesian maps over the full N x N grid (lpRgpu.cu:328)
    size_t NN = (size_t)N * N;
    std::vector<float2> cmapC(3 * NN, make_float2(nanv, nanv));
    for (int t = 0; t < Nc; ++t) {
        int id = cids[t];
        if (id < 0 || (size_t)id >= NN) return h->fail(LP_ERR_ARG, "adj cids index out of range");
        for (int k = 0; k < 3; ++k) cmapC[k * NN + id] = make_float2(C1[k][t], C2[k][t]);
    }
    for (auto& v : cmap) if (!valid2(v)) v = make_float2(nanv, nanv);
    for (auto& v : cmapC) if (!valid2(v)) v = make_float2(nanv, nanv);
    std::vector<float2> fzT = transpose_fz(fz, Nr, Nt);
    int rc;
    std::vector<int2> rng = row_ranges(cmap, Nr, Nt);
    std::vector<lpk::GTile> tA, tC;
    try {
        for (int k = 0; k < 3; ++k) {
            auto coord = [&](int, int r, int c) { return cmap[(size_t)k * npts + (size_t)r * Nt + c]; };
            auto keep = [&](int r0, int c0, int nr, int nc) {
                for (int r = r0; r < r0 + nr; ++r) {
                    int2 q = rng[(size_t)k * Nr + r];
                    if (q.x < c0 + nc && q.y > c0) return true;
                }
                return false;
            };
            auto t = tile_grid(coord, {k}, 0, Nr, Nt, 64, 32, h->box_cap, (h->box_rule >> 2) & 1, keep, h->boxA);
            tA.insert(tA.end(), t.begin(), t.end());
        }
        // LP -> Cartesian: user image tiles, the three spans of a tile consecutive
        const int o = N / 2 - h->N0 / 2;
        auto coordC = [&](int k, int y, int x) { return cmapC[(size_t)k * NN + (size_t)(y + o) * N + x + o]; };
        auto keepC = [](int, int, int, int) { return true; };
        tC = tile_grid(coordC, {0, 1, 2}, 0, h->N0, h->N0, 64, 32, h->box_cap, (h->box_rule >> 3) & 1, keepC, h->boxC);
    } catch (const std::exception& e) {
        return h->fail(LP_ERR_UNSUPPORTED, e.what());
    }
    h->ntilesA = (int)tA.size();
    h->ntilesC = (int)tC.size();
    std::vector<int2> wrng = box_row_ranges(tC, Nr, Nt);
    {   // colour classes of the tiles for the transposed (scatter) use
        const int o = N / 2 - h->N0 / 2;
        auto coordA = [&](int k, int r, int c) { return cmap[(size_t)k * npts + (size_t)r * Nt + c]; };
        auto coordC = [&](int k, int y, int x) { return cmapC[(size_t)k * NN + (size_t)(y + o) * N + x + o]; };
        if ((rc = colour_list(h, tA, coordA, &h->d_permA, &h->d_clsA)) ||
            (rc = colour_list(h, tC, coordC, &h->d_permC, &h->d_clsC)))
            return rc;
    }
    if ((rc = h->upload(&h->d_wrangeA, wrng))) return rc;
    {
        std::vector<int2> in_rows = data_rows(rng, Nr);
        std::vector<unsigned char> out_rows = read_rows(wrng);
        if (!in_rows.empty() && (rc = h->upload(&h->d_inrowsA, in_rows))) return rc;
        if (!out_rows.empty() && (rc = h->upload(&h->d_outrowsA, out_rows))) return rc;
    }
    if ((rc = h->upload(&h->d_tilesA, tA)) || (rc = h->upload(&h->d_tilesC, tC))) return rc;
    if ((rc = stage_tables(h, tA, N, h->P, false, h->stA)) || (rc = stage_tables(h, tC, Nt, Nr, true, h->stC)))
        return rc;
    report_tiles("sinogram->log-polar", tA, h->stA.total);
    report_tiles("log-polar->Cartesian", tC, h->stC.total);
    if ((rc = h->upload(&h->d_cmapA, cmap)) || (rc = h->upload(&h->d_rangeA, rng)) || (rc = h->upload(&h->d_cmapC, cmapC)) ||
        (rc = h->upload(&h->d_fzA, fzT)) || (rc = h->alloc(&h->d_Rc, (size_t)h->Ns * h->P * N)))
        return rc;
    if (any_extra && ((rc = h->upload(&h->d_xrow, xrow)) || (rc = h->upload(&h->d_xent, xent)))) return rc;
    h->nxent = (int)xent.size();
    if (filt) {
        // kernel of the reference filter: h4 = IFFT_4N(ifftshift(filter)) / 4N; keep lags |l| < N.
        // Only those lags reach the N outputs, so any circular length M2 >= 2N - 1 gives the same
        // result; M2 = 2N (padding to a power of two measured slower: twice the points per line).
        int M4 = 4 * N, M2 = 2 * N;
        std::vector<std::complex<double>> a(M4);
        for (int k = 0; k < M4; ++k) a[k] = (double)filt[(k + M4 / 2) % M4];
        lpfft::Plan inv4(M4, +1);
        inv4.exec(a);
        std::vector<std::complex<double>> b(M2, 0.0);
        for (int l = -(N - 1); l <= N - 1; ++l) b[(l + M2) % M2] = a[(l + M4) % M4].real() / M4;
        lpfft::Plan fwd2(M2, -1);
        fwd2.exec(b);
        const int Lh = M2 / 2;
        std::vector<float> H2(Lh + 1);
        for (int k = 0; k <= Lh; ++k) H2[k] = (float)(b[k].real() / (double)M2);
        try {
            h->plan_F = make_plan(Lh);
            h->plan_2N = make_plan(2 * N);
        } catch (const std::exception& e) {
            return h->fail(LP_ERR_UNSUPPORTED, e.what());
        }
        h->Lf = Lh;
        if ((rc = h->upload(&h->d_H2, H2)) || (rc = h->twiddles(Lh)) || (rc = h->twiddles(M2)) ||
            (rc = h->twiddles(2 * N)))
            return rc;
        h->has_filter = true;
    }
    h->has_adj = true;
    if (h->has_filter) h->rtc_tried = false;   // (re)build with the filtered pair pass
    rtc_build(h);
    return LP_OK;
}

// Launch a tiled scatter; with the staged flush (when enabled and built) the
// boxes go to the half-spectrum buffer -- free during every scatter of the
// transposes -- in chunks of slices that fit it, and launch_box_sum forms the
// destination, which needs no clearing; otherwise the destination is cleared
// and the boxes are added with atomics.
static void scatter_to(lp_handle* h, lpk::Scatter g, const lpk::GTile* tiles, int ntiles, int kind,
                       const lp_handle::Stage& st, size_t dst_count, hipStream_t s) {
    const size_t cap = 3 * (size_t)h->Ns * h->h * h->Nr * 2;   // floats of S
    const int chunk = st.ok && st.total > 0 ? (int)std::min<size_t>(cap / (size_t)st.total, (size_t)g.nz) : 0;
    if (h->staged && chunk >= 1) {
        // slices in chunks whose boxes fit the staging area
        const int nz = g.nz;
        for (int z0 = 0; z0 < nz; z0 += chunk) {
            lpk::Scatter c = g;
            c.nz = std::min(chunk, nz - z0);
            c.val = g.val + (size_t)z0 * g.v_z;
            c.dst = g.dst + (size_t)z0 * g.dst_z;
            c.stage = reinterpret_cast<float*>(h->d_S);
            c.stage_z = st.total;
            c.box_off = st.off;
            lpk::launch_scatter(c, tiles, ntiles, kind, s);
            lpk::BoxSum b{};
            b.stage = c.stage; b.stage_z = st.total;
            b.tile_beg = st.tbeg; b.tile_box = st.tbox; b.box_rect = st.rect; b.box_off = st.off;
            b.dst = c.dst; b.dst_z = g.dst_z; b.dst_plane = g.dst_span;
            b.W = g.W; b.H = g.H; b.tiles_x = st.tiles_x; b.tiles_y = st.tiles_y; b.nplanes = st.nplanes;
            b.nz = c.nz;
            lpk::launch_box_sum(b, s);
        }
        return;
    }
    lpk::launch_zero(g.dst, dst_count, s);
    g.stage = nullptr;
    lpk::launch_scatter(g, tiles, ntiles, kind, s);
}

static lpk::RowPass row_pass(lp_handle* h, int nz) {
    lpk::RowPass rp;
    rp.Ntheta = h->Nt; rp.Nrho = h->Nr; rp.h = h->h; rp.R = h->R;
    rp.ldm = row_ldm(h, h->R);
    rp.nz = nz;
    rp.zgroup = std::max(1, std::min(h->zgroup, nz));
    rp.planh = h->plan_h;
    rp.twh = h->tw[h->h];
    rp.twN = h->tw[h->N];
    rp.S = h->d_S;
    rp.conj = 0;
    rp.fft16 = h->fft16;
    rp.in_rows = nullptr;
    rp.out_rows = nullptr;
    const bool rtc = h->rtc_mode && h->rtc_mod;
    rp.rtc_rows_fwd = rtc ? h->rtc_fn[0] : nullptr;
    rp.rtc_rows_inv = rtc ? h->rtc_fn[1] : nullptr;
    rp.rtc_cols = rtc ? h->rtc_fn[2] : nullptr;
    rp.rtc_R = h->rtc_R;
    rp.rtc_rows_nt = h->rtc_rows_nt;
    rp.rtc_cols_nt = h->rtc_cols_nt;
    return rp;
}

// Device buffers of one fwd call: prefiltered image fc [ns0][N][N], the
// row-sweep output tmp (may alias G), the log-polar grids G [3][ns0][Nr][Nt]
// and the half spectra S [3][ns0][h][Nr].
struct FwdScratch {
    float* fc;
    float* tmp;
    float* G;
    float2* S;
};

static int fwd_chain(lp_handle* h, float* R, const float* f, int ns0, hipStream_t s, const FwdScratch& w,
                     bool timed) {
    (void)hipGetLastError();
    const int N = h->N;
    hipEvent_t* ev = timed ? h->prof_begin(0) : nullptr;
    auto mark = [&](int i) { if (ev) (void)hipEventRecord(ev[i], s); };
    mark(0);
    lpk::launch_image_rows(f, h->cubic ? w.tmp : w.fc, ns0, h->N0, N, h->cubic, s);
    mark(1);
    if (h->stop_after == 1) return check_launch(h, "lp_fwd launch");
    if (h->cubic) lpk::launch_col_iir(w.tmp, w.fc, ns0, N, N, h->colchunk, 1, s);
    mark(2);
    if (h->stop_after == 2) return check_launch(h, "lp_fwd launch");
    lpk::RowPass rp = row_pass(h, ns0);
    rp.S = w.S;
    if (!h->stop_after) {   // (stage-wise tests read every row)
        rp.in_rows = h->d_inrowsF;
        rp.out_rows = h->d_outrowsF;
    }
    const long long grid = (long long)h->Nr * h->Nt;
    {   // image -> log-polar samples x e^rho into F (aliases G)
        lpk::Gather g{};
        g.src = w.fc; g.src_z = (long long)N * N; g.src_span = 0; g.W = N; g.H = N;
        g.coord = h->d_cmapF; g.c_span = grid; g.c_row = h->Nt; g.c_ro = 0; g.c_co = 0;
        g.scale = h->d_erho; g.s_row = h->Nt;
        g.out = w.G; g.o_z = grid; g.o_span = (long long)ns0 * grid; g.o_row = h->Nt;
        g.nz = ns0; g.zgroup = h->gzgroup; g.group = 1; g.cubic = h->cubic; g.box_floats = h->boxF;
        if (h->gather_direct) lpk::launch_gather_direct(g, h->d_rangeF, h->Nr, h->Nt, s);
        else lpk::launch_gather(g, h->d_tilesF, h->ntilesF, 0, s);
    }
    mark(3);
    if (h->stop_after == 3) return check_launch(h, "lp_fwd launch");
    lpk::launch_rows_fwd(rp, w.G, h->d_rangeF, s);
    mark(4);
    if (h->stop_after == 4) return check_launch(h, "lp_fwd launch");
    lpk::launch_cols(rp, h->plan_M, h->tw[h->Nr], h->d_fzF, s);
    mark(5);
    if (h->stop_after == 5) return check_launch(h, "lp_fwd launch");
    if (!h->stop_after && h->fuse_polar && h->R == kFuseBase + 3) {
        // inverse row pass and polar gather in one kernel (G stays on chip)
        lpk::PolarOut po;
        po.out = R; po.o_z = (long long)h->P0 * h->N0;
        po.coord = h->d_pcoord; po.off = h->d_poff; po.band_beg = h->d_pband; po.nband = h->npband;
        po.cubic = h->cubic;
        if (lpk::launch_rows_inv_polar(rp, po, s)) {
            mark(6);
            mark(7);
            return check_launch(h, "lp_fwd launch");
        }
    }
    // only the columns the polar gather reads (all of them when debugging stage 6)
    lpk::launch_rows_inv(rp, w.G, h->stop_after ? nullptr : h->d_wrangeF, s);
    mark(6);
    if (h->stop_after == 6) return check_launch(h, "lp_fwd launch");
    {   // log-polar -> polar for the kept angles, cropped detector
        lpk::Gather g{};
        g.src = w.G; g.src_z = grid; g.src_span = (long long)ns0 * grid; g.W = h->Nt; g.H = h->Nr;
        g.coord = h->d_p2lp; g.c_span = 0; g.c_row = h->os * N; g.c_ro = 0; g.c_co = N / 2 - h->cor;
        g.scale = nullptr; g.s_row = 0;
        g.out = R; g.o_z = (long long)h->P0 * h->N0; g.o_span = 0; g.o_row = h->N0;
        g.nz = ns0; g.zgroup = h->pzgroup; g.group = 1; g.cubic = h->cubic; g.box_floats = h->boxP;
        lpk::launch_gather(g, h->d_tilesP, h->ntilesP, 2, s);
    }
    mark(7);
    return check_launch(h, "lp_fwd launch");
}

int lp_fwd(lp_handle* h, float* R, const float* f, int ns0, void* stream) {
    if (!h) return LP_ERR_ARG;
    if (!h->has_fwd) return h->fail(LP_ERR_ARG, "forward tables not initialised (lp_init_fwd)");
    if (!R || !f) return h->fail(LP_ERR_ARG, "null data pointer");
    if (ns0 < 1 || ns0 > h->Ns) return h->fail(LP_ERR_ARG, "ns0 must be in [1, Nslices]");
    DeviceGuard guard(h->device);
    return fwd_chain(h, R, f, ns0, (hipStream_t)stream, FwdScratch{h->d_fc, h->d_tmp, h->d_G, h->d_S}, true);
}

int lp_fwd_noalloc(lp_handle* h, float* R, const float* f, int ns0, void* stream) {
    if (!h) return LP_ERR_ARG;
    if (!h->has_fwd) return h->fail(LP_ERR_ARG, "forward tables not initialised (lp_init_fwd)");
    if (!R || !f) return h->fail(LP_ERR_ARG, "null data pointer");
    if (ns0 < 1 || ns0 > h->Ns) return h->fail(LP_ERR_ARG, "ns0 must be in [1, Nslices]");
    const size_t grid = (size_t)h->Nr * h->Nt;
    const size_t b_fc = round256((size_t)ns0 * h->N * h->N * sizeof(float));
    const size_t b_G = round256(3 * (size_t)ns0 * grid * sizeof(float));
    const size_t b_S = round256(3 * (size_t)ns0 * h->h * h->Nr * sizeof(float2));
    // the row-sweep output (N x N per slice) reuses G, which is 6 N^2 per slice
    if (b_fc + b_G + b_S > kNoAllocScratchBytes)
        return h->fail(LP_ERR_ARG, "lp_fwd_noalloc: " + std::to_string(ns0) + " slice(s) at N = " +
                                       std::to_string(h->N) + " need " + std::to_string(b_fc + b_G + b_S) +
                                       " bytes of scratch; the static scratch holds " +
                                       std::to_string(kNoAllocScratchBytes) + " bytes");
    DeviceGuard guard(h->device);
    if (!guard.ok) return h->fail(LP_ERR_DEVICE, "hipSetDevice failed");
    void* base = nullptr;
    if (int rc = scratch_address(h->device, &base)) return h->fail(rc, "hipGetSymbolAddress(scratch) failed");
    unsigned char* p = static_cast<unsigned char*>(base);
    float* fc = reinterpret_cast<float*>(p);
    float* G = reinterpret_cast<float*>(p + b_fc);
    float2* S = reinterpret_cast<float2*>(p + b_fc + b_G);
    return fwd_chain(h, R, f, ns0, (hipStream_t)stream, FwdScratch{fc, G, G, S}, false);
}

int lp_adj(lp_handle* h, float* f, const float* R, int ns0, void* stream) {
    if (!h) return LP_ERR_ARG;
    if (!h->has_adj) return h->fail(LP_ERR_ARG, "adjoint tables not initialised (lp_init_adj)");
    if (!R || !f) return h->fail(LP_ERR_ARG, "null data pointer");
    if (ns0 < 1 || ns0 > h->Ns) return h->fail(LP_ERR_ARG, "ns0 must be in [1, Nslices]");
    DeviceGuard guard(h->device);
    hipStream_t s = (hipStream_t)stream;
    (void)hipGetLastError();
    const int N = h->N;
    lpk::SinoPrep sp;
    sp.N0 = h->N0; sp.N = N; sp.P0 = h->P0; sp.P = h->P; sp.os = h->os; sp.lb = N / 2 - h->cor;
    sp.H2 = h->has_filter ? h->d_H2 : nullptr;
    sp.Lh = h->has_filter ? h->Lf : N;
    sp.planN = h->has_filter ? h->plan_F : h->plan_N;
    sp.twN = h->has_filter ? h->tw[h->Lf] : nullptr;
    sp.tw2N = h->has_filter ? h->tw[2 * h->Lf] : nullptr;
    sp.plan2N = h->plan_2N;
    sp.tw2Nc = h->has_filter ? h->tw[2 * N] : nullptr;
    sp.rtc_pair = h->rtc_mode && h->rtc_mod ? h->rtc_fn[3] : nullptr;
    sp.rtc_pair_nt = h->rtc_pair_nt;
    hipEvent_t* ev = h->prof_begin(1);
    auto mark = [&](int i) { if (ev) (void)hipEventRecord(ev[i], s); };
    mark(0);
    lpk::launch_sino_rows(R, h->cubic ? h->d_tmp : h->d_Rc, ns0, sp, h->cubic, s);
    mark(1);
    if (h->stop_after == 1) return check_launch(h, "lp_adj launch");
    if (h->cubic) lpk::launch_col_iir(h->d_tmp, h->d_Rc, ns0, h->P, N, h->colchunk, h->os, s);
    mark(2);
    if (h->stop_after == 2) return check_launch(h, "lp_adj launch");
    lpk::RowPass rp = row_pass(h, ns0);
    if (!h->stop_after) {   // (stage-wise tests read every row)
        rp.in_rows = h->d_inrowsA;
        rp.out_rows = h->d_outrowsA;
    }
    lpk::Src src{h->d_Rc, N, h->P, (long long)h->P * N};
    const long long grid = (long long)h->Nr * h->Nt;
    {   // sinogram -> log-polar samples into F (aliases G)
        lpk::Gather g{};
        g.src = h->d_Rc; g.src_z = (long long)h->P * N; g.src_span = 0; g.W = N; g.H = h->P;
        g.coord = h->d_cmapA; g.c_span = grid; g.c_row = h->Nt; g.c_ro = 0; g.c_co = 0;
        g.scale = nullptr; g.s_row = 0;
        g.out = h->d_G; g.o_z = grid; g.o_span = (long long)ns0 * grid; g.o_row = h->Nt;
        g.nz = ns0; g.zgroup = h->gzgroup; g.group = 1; g.cubic = h->cubic; g.box_floats = h->boxA;
        lpk::launch_gather(g, h->d_tilesA, h->ntilesA, 1, s);
        // samples that share a log-polar point with another one
        lpk::launch_add_extras(h->d_G, ns0, h->Nt, h->Nr, src, h->d_xrow, h->d_xent, h->nxent, h->cubic, s);
    }
    mark(3);
    if (h->stop_after == 3) return check_launch(h, "lp_adj launch");
    lpk::launch_rows_fwd(rp, h->d_G, h->d_rangeA, s);
    mark(4);
    if (h->stop_after == 4) return check_launch(h, "lp_adj launch");
    lpk::launch_cols(rp, h->plan_M, h->tw[h->Nr], h->d_fzA, s);
    mark(5);
    if (h->stop_after == 5) return check_launch(h, "lp_adj launch");
    lpk::launch_rows_inv(rp, h->d_G, h->stop_after ? nullptr : h->d_wrangeA, s);
    mark(6);
    if (h->stop_after == 6) return check_launch(h, "lp_adj launch");
    {   // log-polar -> Cartesian, three spans summed, cropped
        const int o = N / 2 - h->N0 / 2;
        lpk::Gather g{};
        g.src = h->d_G; g.src_z = grid; g.src_span = (long long)ns0 * grid; g.W = h->Nt; g.H = h->Nr;
        g.coord = h->d_cmapC; g.c_span = (long long)N * N; g.c_row = N; g.c_ro = o; g.c_co = o;
        g.scale = nullptr; g.s_row = 0;
        g.out = f; g.o_z = (long long)h->N0 * h->N0; g.o_span = 0; g.o_row = h->N0;
        g.nz = ns0; g.zgroup = h->gzgroup; g.group = 3; g.cubic = h->cubic; g.box_floats = h->boxC;
        lpk::launch_gather(g, h->d_tilesC, h->ntilesC, 3, s);
    }
    mark(7);
    return check_launch(h, "lp_adj launch");
}

// Exact transpose of lp_fwd: R [ns0][P0][N0] -> f [ns0][N0][N0].  The stages
// of lp_fwd in reverse order, each replaced by its transpose.
int lp_fwd_T(lp_handle* h, float* f, const float* R, int ns0, void* stream) {
    if (!h) return LP_ERR_ARG;
    if (!h->has_fwd) return h->fail(LP_ERR_ARG, "forward tables not initialised (lp_init_fwd)");
    if (!R || !f) return h->fail(LP_ERR_ARG, "null data pointer");
    if (ns0 < 1 || ns0 > h->Ns) return h->fail(LP_ERR_ARG, "ns0 must be in [1, Nslices]");
    DeviceGuard guard(h->device);
    hipStream_t s = (hipStream_t)stream;
    (void)hipGetLastError();
    const int N = h->N;
    const long long grid = (long long)h->Nr * h->Nt;
    const int2* g_range = nullptr;   // theta columns of the transposed grid that can be non-zero
    if (h->pt_ok && h->polar_t) {
        // transpose of the log-polar -> polar gather as two gathers; the
        // intermediate lives in the half-spectrum buffer, free until the row pass
        lpk::PolarT p{};
        p.val = R; p.v_z = (long long)h->P0 * h->N0;
        p.coord = h->d_p2lp; p.c_row = h->os * N; p.c_co = N / 2 - h->cor;
        p.cxa = h->d_pt_cxa; p.u = reinterpret_cast<float*>(h->d_S); p.nA = h->pt_nA; p.abeg = h->d_pt_abeg;
        p.xs = h->d_pt_xs; p.qlo = h->pt_qlo; p.Q = h->pt_Q;
        p.G = h->d_G; p.aj = h->d_pt_aj; p.jlo = h->pt_jlo; p.jn = h->pt_jn;
        p.Nr = h->Nr; p.Nt = h->Nt; p.N0 = h->N0; p.nz = ns0; p.cubic = h->cubic;
        lpk::launch_polar_transpose(p, s);
        g_range = h->d_wrangeF;   // the polar taps' columns; zero elsewhere (not written)
    } else {
        lpk::launch_zero(h->d_G, 3 * (size_t)ns0 * grid, s);
        // transpose of the log-polar -> polar gather
        lpk::Scatter g{};
        g.dst = h->d_G; g.dst_z = grid; g.dst_span = (long long)ns0 * grid; g.W = h->Nt; g.H = h->Nr;
        g.coord = h->d_p2lp; g.c_span = 0; g.c_row = h->os * N; g.c_ro = 0; g.c_co = N / 2 - h->cor;
        g.scale = nullptr; g.s_row = 0;
        g.val = R; g.v_z = (long long)h->P0 * h->N0; g.v_span = 0; g.v_row = h->N0;
        g.nz = ns0; g.zgroup = h->gzgroup; g.group = 1; g.cubic = h->cubic; g.box_floats = h->boxP;
        g.perm = h->d_permP; g.cls = h->d_clsP;
        lpk::launch_scatter(g, h->d_tilesP, h->ntilesP, 2, s);
    }
    lpk::RowPass rp = row_pass(h, ns0);
    rp.conj = 1;
    lpk::Src src{h->d_fc, N, N, (long long)N * N};
    lpk::launch_rows_fwd(rp, h->d_G, g_range, s);
    lpk::launch_cols(rp, h->plan_M, h->tw[h->Nr], h->d_fzF, s);
    lpk::launch_rows_inv(rp, h->d_G, h->d_rangeF, s);   // the scatter reads the sampled columns only
    float* acc = h->d_fc;   // never aliases G, which the scatter reads
    {   // transpose of the image -> log-polar gather (x e^rho)
        lpk::Scatter g{};
        g.dst = acc; g.dst_z = (long long)N * N; g.dst_span = 0; g.W = N; g.H = N;
        g.coord = h->d_cmapF; g.c_span = grid; g.c_row = h->Nt; g.c_ro = 0; g.c_co = 0;
        g.scale = h->d_erho; g.s_row = h->Nt;
        g.val = h->d_G; g.v_z = grid; g.v_span = (long long)ns0 * grid; g.v_row = h->Nt;
        g.nz = ns0; g.zgroup = h->gzgroup; g.group = 1; g.cubic = h->cubic; g.box_floats = h->boxF;
        g.perm = h->d_permF; g.cls = h->d_clsF;
        scatter_to(h, g, h->d_tilesF, h->ntilesF, 0, h->stF, (size_t)ns0 * N * N, s);
    }
    if (h->cubic) {
        lpk::launch_col_iir_T(h->d_fc, h->d_tmp, ns0, N, N, h->colchunk, s);
        lpk::launch_image_rows_T(h->d_tmp, f, ns0, h->N0, N, true, s);
    } else {
        lpk::launch_image_rows_T(acc, f, ns0, h->N0, N, false, s);
    }
    return check_launch(h, "lp_fwd_T launch");
}

// Exact transpose of lp_adj: f [ns0][N0][N0] -> R [ns0][P0][N0].
int lp_adj_T(lp_handle* h, float* R, const float* f, int ns0, void* stream) {
    if (!h) return LP_ERR_ARG;
    if (!h->has_adj) return h->fail(LP_ERR_ARG, "adjoint tables not initialised (lp_init_adj)");
    if (!R || !f) return h->fail(LP_ERR_ARG, "null data pointer");
    if (ns0 < 1 || ns0 > h->Ns) return h->fail(LP_ERR_ARG, "ns0 must be in [1, Nslices]");
    DeviceGuard guard(h->device);
    hipStream_t s = (hipStream_t)stream;
    (void)hipGetLastError();
    const int N = h->N;
    const long long grid = (long long)h->Nr * h->Nt;
    {   // transpose of the log-polar -> Cartesian gather (three spans per tile)
        const int o = N / 2 - h->N0 / 2;
        lpk::Scatter g{};
        g.dst = h->d_G; g.dst_z = grid; g.dst_span = (long long)ns0 * grid; g.W = h->Nt; g.H = h->Nr;
        g.coord = h->d_cmapC; g.c_span = (long long)N * N; g.c_row = N; g.c_ro = o; g.c_co = o;
        g.scale = nullptr; g.s_row = 0;
        g.val = f; g.v_z = (long long)h->N0 * h->N0; g.v_span = 0; g.v_row = h->N0;
        g.nz = ns0; g.zgroup = h->gzgroup; g.group = 3; g.cubic = h->cubic; g.box_floats = h->boxC;
        g.perm = h->d_permC; g.cls = h->d_clsC;
        scatter_to(h, g, h->d_tilesC, h->ntilesC, 3, h->stC, 3 * (size_t)ns0 * grid, s);
    }
    lpk::RowPass rp = row_pass(h, ns0);
    rp.conj = 1;
    lpk::Src src{h->d_Rc, N, h->P, (long long)h->P * N};
    lpk::launch_rows_fwd(rp, h->d_G, nullptr, s);
    lpk::launch_cols(rp, h->plan_M, h->tw[h->Nr], h->d_fzA, s);
    lpk::launch_rows_inv(rp, h->d_G, h->d_rangeA, s);
    float* acc = h->d_Rc;
    {   // transpose of the sinogram -> log-polar gather, plus its extra samples
        lpk::Scatter g{};
        g.dst = acc; g.dst_z = (long long)h->P * N; g.dst_span = 0; g.W = N; g.H = h->P;
        g.coord = h->d_cmapA; g.c_span = grid; g.c_row = h->Nt; g.c_ro = 0; g.c_co = 0;
        g.scale = nullptr; g.s_row = 0;
        g.val = h->d_G; g.v_z = grid; g.v_span = (long long)ns0 * grid; g.v_row = h->Nt;
        g.nz = ns0; g.zgroup = h->gzgroup; g.group = 1; g.cubic = h->cubic; g.box_floats = h->boxA;
        g.perm = h->d_permA; g.cls = h->d_clsA;
        scatter_to(h, g, h->d_tilesA, h->ntilesA, 1, h->stA, (size_t)ns0 * h->P * N, s);
        lpk::launch_scatter_extra(h->d_G, acc, ns0, h->Nt, h->Nr, N, h->P, h->d_xrow, h->d_xent, h->cubic, s);
    }
    lpk::SinoPrep sp;
    sp.N0 = h->N0; sp.N = N; sp.P0 = h->P0; sp.P = h->P; sp.os = h->os; sp.lb = N / 2 - h->cor;
    sp.H2 = h->has_filter ? h->d_H2 : nullptr;
    sp.Lh = h->has_filter ? h->Lf : N;
    sp.planN = h->has_filter ? h->plan_F : h->plan_N;
    sp.twN = h->has_filter ? h->tw[h->Lf] : nullptr;
    sp.tw2N = h->has_filter ? h->tw[2 * h->Lf] : nullptr;
    sp.plan2N = h->plan_2N;
    sp.tw2Nc = h->has_filter ? h->tw[2 * N] : nullptr;
    sp.rtc_pair = nullptr;
    sp.rtc_pair_nt = 0;
    if (h->cubic) {
        lpk::launch_col_iir_T(acc, h->d_tmp, ns0, h->P, N, h->colchunk, s);
        lpk::launch_sino_rows_T(h->d_tmp, R, ns0, sp, true, s);
    } else {
        lpk::launch_sino_rows_T(acc, R, ns0, sp, false, s);
    }
    return check_launch(h, "lp_adj_T launch");
}

void lp_destroy(lp_handle* h) { delete h; }

const char* lp_last_error(const lp_handle* h) {
    if (!h) return g_create_err.c_str();
    return h->err.c_str();
}

int lp_stream_bench(int device, float* dst, const float* src, size_t n_floats, int kind, int blocks, void* stream) {
    if (!dst || (kind != 2 && !src) || n_floats % 4 || kind < 0 || kind > 2) return LP_ERR_ARG;
    DeviceGuard guard(device);
    if (!guard.ok) return LP_ERR_DEVICE;
    (void)hipGetLastError();
    if (lpk::launch_stream(kind, dst, src, n_floats / 4, blocks, (hipStream_t)stream)) return LP_ERR_ARG;
    return hipGetLastError() == hipSuccess ? LP_OK : LP_ERR_DEVICE;
}

int lp_mem_info(int device, size_t* free_bytes, size_t* total_bytes) {
    DeviceGuard guard(device);
    if (!guard.ok) return LP_ERR_DEVICE;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return LP_ERR_DEVICE;
    if (free_bytes) *free_bytes = fr;
    if (total_bytes) *total_bytes = tot;
    return LP_OK;
}

int lp_info(const lp_handle* h, int64_t* info) {
    if (!h || !info) return LP_ERR_ARG;
    int64_t v[12] = {h->N, h->N0, h->Nt, h->Nr, h->P, h->P0, h->os, h->Ns, h->cor, h->cubic ? 1 : 0,
                     h->has_filter ? 1 : 0, (int64_t)h->bytes};
    std::memcpy(info, v, sizeof(v));
    return LP_OK;
}

int lp_set_tuning(lp_handle* h, int key, int value) {
    if (!h) return LP_ERR_ARG;
    switch (key) {
        case 0: {
            if (value < 1 || value > 16 || (value & (value - 1))) return h->fail(LP_ERR_ARG, "rows per block must be a power of two in [1, 16]");
            int old = h->R;
            h->R = value;
            int r = pick_rows(h);
            if (r != value) { h->R = old; return h->fail(LP_ERR_ARG, "rows per block exceeds the LDS or FFT budget"); }
            return LP_OK;
        }
        case 1:
            if (value < 1) return h->fail(LP_ERR_ARG, "slices per block must be >= 1");
            h->zgroup = value;
            return LP_OK;
        case 2:
            if (value != 32 && value != 64 && value != 128) return h->fail(LP_ERR_ARG, "column chunk must be 32, 64 or 128");
            h->colchunk = value;
            return LP_OK;
        case 3:
            if (value < 1) return h->fail(LP_ERR_ARG, "slices per gather block must be >= 1");
            h->gzgroup = value;
            return LP_OK;
        case 4:
            if (value < 0 || value > 2) return h->fail(LP_ERR_ARG, "FFT kernel family must be 0, 1 or 2");
            h->fft16 = value;
            return LP_OK;
        case 6:
            if (value != 0 && value != 1) return h->fail(LP_ERR_ARG, "polar fusion must be 0 or 1");
            h->fuse_polar = value;
            return LP_OK;
        case 7:
            if (value != 0 && value != 1) return h->fail(LP_ERR_ARG, "separable polar transpose must be 0 or 1");
            h->polar_t = value;
            return LP_OK;
        case 8:
            if (value != 0 && value != 1) return h->fail(LP_ERR_ARG, "staged scatter flush must be 0 or 1");
            h->staged = value;
            return LP_OK;
        case 5:
            if (h->has_fwd || h->has_adj)
                return h->fail(LP_ERR_ARG, "the gather box budget is fixed once lp_init_fwd / lp_init_adj ran");
            if (value < 1024 || value > 32768) return h->fail(LP_ERR_ARG, "gather box budget must be in [1024, 32768] floats");
            h->box_cap = value;
            return LP_OK;
        case 11:
            if (value != 0 && value != 256 && value != 512 && value != 1024)
                return h->fail(LP_ERR_ARG, "threads of the compiled column pass must be 0, 256, 512 or 1024");
            if ((long long)value * 32 < h->Nr && value)
                return h->fail(LP_ERR_ARG, "too few threads for the column length");
            h->rtc_cols_req = value;
            h->rtc_tried = false;
            if (h->rtc_mode && (h->has_fwd || h->has_adj)) {
                DeviceGuard guard(h->device);
                rtc_build(h);
            }
            return LP_OK;
        case 14:
            if (value != 0 && value != 1) return h->fail(LP_ERR_ARG, "direct gather must be 0 or 1");
            h->gather_direct = value;
            return LP_OK;
        case 13:
            if (value < 1) return h->fail(LP_ERR_ARG, "slices per polar gather block must be >= 1");
            h->pzgroup = value;
            return LP_OK;
        case 12:
            if (value < 0 || value > 15) return h->fail(LP_ERR_ARG, "compiled pass variant must be in [0, 15]");
            h->rtc_cols_var = value;
            h->rtc_tried = false;
            if (h->rtc_mode && (h->has_fwd || h->has_adj)) {
                DeviceGuard guard(h->device);
                rtc_build(h);
            }
            return LP_OK;
        case 10:
            if (value != 0 && value != 1) return h->fail(LP_ERR_ARG, "run-time compiled passes must be 0 or 1");
            h->rtc_mode = value;
            if (value && (h->has_fwd || h->has_adj)) {
                DeviceGuard guard(h->device);
                rtc_build(h);
            }
            return LP_OK;
        case 9:
            if (h->has_fwd || h->has_adj)
                return h->fail(LP_ERR_ARG, "the gather box stride is fixed once lp_init_fwd / lp_init_adj ran");
            if (value < 0 || value > 15) return h->fail(LP_ERR_ARG, "box stride rule must be a 4-bit mask");
            h->box_rule = value;
            return LP_OK;
    }
    return h->fail(LP_ERR_ARG, "unknown tuning key");
}

int lp_profile(lp_handle* h, int enable) {
    if (!h) return LP_ERR_ARG;
    DeviceGuard guard(h->device);
    h->prof_clear();
    h->prof = enable != 0;
    return LP_OK;
}

int lp_profile_read(lp_handle* h, int which, float* ms, int* ncalls) {
    if (!h || !ms || (which != 0 && which != 1)) return LP_ERR_ARG;
    DeviceGuard guard(h->device);
    for (int i = 0; i < lp_handle::kNStage; ++i) ms[i] = 0.f;
    size_t n = h->prof_used[which];
    for (size_t c = 0; c < n; ++c) {
        auto& ev = h->prof_sets[which][c];
        if (hipEventSynchronize(ev[lp_handle::kNStage]) != hipSuccess)
            return h->fail(LP_ERR_DEVICE, "hipEventSynchronize failed");
        for (int i = 0; i < lp_handle::kNStage; ++i) {
            float t = 0.f;
            if (hipEventElapsedTime(&t, ev[i], ev[i + 1]) != hipSuccess)
                return h->fail(LP_ERR_DEVICE, "hipEventElapsedTime failed");
            ms[i] += t;
        }
    }
    if (ncalls) *ncalls = (int)n;
    return LP_OK;
}

int lp_debug_stop(lp_handle* h, int stage) {
    if (!h) return LP_ERR_ARG;
    if (stage < 0 || stage > 6) return h->fail(LP_ERR_ARG, "stage must be in [0, 6]");
    h->stop_after = stage;
    return LP_OK;
}

int lp_debug_copy(lp_handle* h, int which, void* dst, size_t bytes) {
    if (!h || !dst) return LP_ERR_ARG;
    DeviceGuard guard(h->device);
    const void* src = nullptr;
    size_t cap = 0;
    switch (which) {
        case 0: src = h->d_fc; cap = (size_t)h->Ns * h->N * h->N * 4; break;
        case 1: src = h->d_Rc; cap = (size_t)h->Ns * h->P * h->N * 4; break;
        case 2: src = h->d_G; cap = 3 * (size_t)h->Ns * h->Nr * h->Nt * 4; break;
        case 3: src = h->d_S; cap = 3 * (size_t)h->Ns * h->h * h->Nr * 8; break;
        case 4: src = h->d_tmp; cap = (size_t)h->Ns * std::max((size_t)h->N * h->N, (size_t)h->P * h->N) * 4; break;
        default: return h->fail(LP_ERR_ARG, "unknown buffer");
    }
    if (!src) return h->fail(LP_ERR_ARG, "buffer not allocated (init_fwd / init_adj)");
    if (bytes > cap) return h->fail(LP_ERR_ARG, "copy larger than the buffer");
    if (hipDeviceSynchronize() != hipSuccess || hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) != hipSuccess)
        return h->fail(LP_ERR_DEVICE, "device to host copy failed");
    return LP_OK;
}

int lp_get_tuning(const lp_handle* h, int key, int* value) {
    if (!h || !value) return LP_ERR_ARG;
    switch (key) {
        case 0: *value = h->R; return LP_OK;
        case 1: *value = h->zgroup; return LP_OK;
        case 2: *value = h->colchunk; return LP_OK;
        case 3: *value = h->gzgroup; return LP_OK;
        case 4: *value = h->fft16; return LP_OK;
        case 5: *value = h->box_cap; return LP_OK;
        case 6: *value = h->fuse_polar; return LP_OK;
        case 7: *value = h->pt_ok ? h->polar_t : 0; return LP_OK;
        case 8: *value = h->staged; return LP_OK;
        case 9: *value = h->box_rule; return LP_OK;
        case 10: *value = h->rtc_mode && h->rtc_mod ? 1 : 0; return LP_OK;
        case 11: *value = h->rtc_cols_nt; return LP_OK;
        case 12: *value = h->rtc_cols_var; return LP_OK;
        case 13: *value = h->pzgroup; return LP_OK;
        case 14: *value = h->gather_direct; return LP_OK;
    }
    return LP_ERR_ARG;
}

}  // extern "C"
