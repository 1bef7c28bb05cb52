// Host precompute of the log-polar operator (FP64 math, FP32 storage).
//
// Produces the reference's flat parameter arrays, byte layout identical to
// what src/cuda/params.cu:40-145 parses:
//   glpars      src/lprec/initsgl.py:25-57
//   fwdparsi/f  src/lprec/initsfwd.py:15-66, 96-126
//   adjparsi/f  src/lprec/initsadj.py:19-116, 205-244
// The reference builds these with cupy; here every quantity is evaluated in
// double precision except the float32 log-polar lattice (thsp/rhosp,
// initsgl.py:42-43) and the float32 span-membership comparisons that the
// reference's dtype rules imply (initsadj.py:46, 57, 64).  Compiled with
// -ffp-contract=off so that the arithmetic is plain IEEE double.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "fft64.hpp"
#include "lp_internal.hpp"

using lpfft::cd;

namespace {

thread_local std::string g_pre_err;

std::vector<double> linspace(double a, double b, int n) {
    // numpy.linspace: start + arange(n) * step, last element set to stop
    std::vector<double> v(n);
    if (n == 1) { v[0] = a; return v; }
    double step = (b - a) / (double)(n - 1);
    for (int j = 0; j < n; ++j) v[j] = (double)j * step + a;
    v[n - 1] = b;
    return v;
}

// initsgl.py:80-99 followed by fft(ifftshift(.)) (initsgl.py:46-49): the
// 5 non-zero taps sit at the centre, so the transform is evaluated directly.
std::vector<cd> spline_b3_hat(const std::vector<float>& grid) {
    int n = (int)grid.size();
    std::vector<double> x2(n);
    double mid = ((double)grid[n - 1] + (double)grid[0]) / 2;
    for (int j = 0; j < n; ++j) x2[j] = (double)grid[j] - mid;
    double stepx = x2[1] - x2[0];
    int ri = 2;
    double r = 1.0 * stepx;
    int c = (int)std::ceil((n + 1) / 2.0);
    double x2c = x2[c - 1];
    std::vector<double> b3(2 * ri + 1, 0.0);
    for (int ix = -ri; ix <= ri; ++ix) {
        double d = std::fabs(x2[c - 1 + ix] - x2c) / r;
        double v = 0.0;
        if (d < 1) v = (3 * d * d * d - 6 * d * d + 4) / 6;
        else if (d < 2) v = (-d * d * d + 6 * d * d - 12 * d + 8) / 6;
        b3[ix + ri] = v;
    }
    // element at centre index c-1 moves to index 0 under ifftshift (n even:
    // shift by -n/2, and c-1 == n/2)
    int centre = c - 1;
    int shift = n / 2;
    std::vector<cd> out(n);
    for (int k = 0; k < n; ++k) {
        cd acc(0, 0);
        for (int ix = -ri; ix <= ri; ++ix) {
            int pos = ((centre + ix - shift) % n + n) % n;
            double a = -2.0 * M_PI * (double)((long long)k * pos % n) / (double)n;
            acc += b3[ix + ri] * cd(std::cos(a), std::sin(a));
        }
        out[k] = acc;
    }
    return out;
}

// initsfwd.py:74-83 / initsadj.py:125-134
std::vector<double> h_weights(int nl) {
    static const double num[11] = {-216254335.0, 679543284.0, -1412947389.0, 2415881496.0,
                                   -3103579086.0, 2939942400.0, -2023224114.0, 984515304.0,
                                   -321455811.0, 63253516.0, -5675265.0};
    double corr[11];
    for (int j = 0; j < 11; ++j) corr[j] = 1 + num[j] / 958003200.0;
    corr[0] = 2 * (corr[0] - 0.5);
    std::vector<double> h(nl, 1.0);
    h[0] *= corr[0];
    for (int j = 1; j < 11; ++j) {
        h[j] *= corr[j];
        h[nl - j] *= corr[j];
    }
    return h;
}

template <class F>
void parallel_rows(int n, F&& fn) {
    unsigned hw = std::thread::hardware_concurrency();
    int nt = (int)std::max(1u, std::min(hw ? hw : 1u, 16u));
    nt = std::min(nt, n);
    if (nt <= 1) { fn(0, n); return; }
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) {
        int lo = (int)((long long)n * t / nt), hi = (int)((long long)n * (t + 1) / nt);
        th.emplace_back([&fn, lo, hi] { fn(lo, hi); });
    }
    for (auto& t : th) t.join();
}

// fzeta_loop_weights (initsfwd.py:69-93) / fzeta_loop_weights_adj
// (initsadj.py:119-143) composed with the outer fftshift and the [:, :H]
// slice of create_fwd/create_adj: returns the [Nrho][H] half spectrum (before
// the constant and the B-spline compensation).
std::vector<cd> fzeta_half(int Ntheta, int Nrho, double betas, double rhos, bool adjoint) {
    int nl = 4 * Ntheta;
    int H = Ntheta / 2 + 1;
    std::vector<double> thl(nl), logcos(nl);
    for (int j = 0; j < nl; ++j) thl[j] = ((double)j - nl / 2) / (double)nl * betas;
    for (int j = 0; j < nl; ++j) logcos[j] = std::log(std::cos(thl[j]));
    std::vector<double> h = h_weights(nl);
    double dth = thl[1] - thl[0];
    std::vector<cd> out((size_t)Nrho * H, cd(0, 0));
    lpfft::Plan plan(nl, -1);
    parallel_rows(Nrho, [&](int lo, int hi) {
        std::vector<cd> v(nl);
        for (int i = lo; i < hi; ++i) {
            // output row i <- fzeta row (i + Nrho/2) mod Nrho, kappa = that - Nrho/2
            int jr = (i + Nrho / 2) % Nrho;
            if (jr == 0) continue;  // fZ[0] = 0
            double kappa = (double)jr - Nrho / 2;
            cd e = adjoint ? cd(0.0, 2 * M_PI * kappa / rhos) : cd(-1.0, -2 * M_PI * kappa / rhos);
            for (int n = 0; n < nl; ++n) {
                int src = (n + nl / 2) % nl;  // fftshift of the input
                cd z = e * logcos[src];
                v[n] = h[src] * std::exp(z);
            }
            plan.exec(v);
            for (int c = 0; c < H; ++c) {
                if (c == Ntheta / 2) continue;  // fZ[:, 0] = 0 (Nyquist column)
                out[(size_t)i * H + c] = v[c] * dth;
            }
        }
    });
    return out;
}

// Gauss-Jordan inverse with partial pivoting (numpy.linalg.inv in the reference).
bool invert(std::vector<double>& a, int n) {
    std::vector<double> inv(n * n, 0.0);
    for (int i = 0; i < n; ++i) inv[i * n + i] = 1.0;
    for (int c = 0; c < n; ++c) {
        int p = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(a[r * n + c]) > std::fabs(a[p * n + c])) p = r;
        if (a[p * n + c] == 0.0) return false;
        if (p != c)
            for (int k = 0; k < n; ++k) { std::swap(a[p * n + k], a[c * n + k]); std::swap(inv[p * n + k], inv[c * n + k]); }
        double d = a[c * n + c];
        for (int k = 0; k < n; ++k) { a[c * n + k] /= d; inv[c * n + k] /= d; }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            double f = a[r * n + c];
            if (f == 0.0) continue;
            for (int k = 0; k < n; ++k) { a[r * n + k] -= f * a[c * n + k]; inv[r * n + k] -= f * inv[c * n + k]; }
        }
    }
    a = inv;
    return true;
}

// wint (initsadj.py:176-203)
std::vector<double> wint(int n, const std::vector<double>& t) {
    int N = (int)t.size();
    std::vector<double> s = linspace(1e-40, 1, n);
    std::vector<double> V(n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) V[i * n + j] = std::exp((double)i * std::log(s[j]));
    invert(V, n);
    // u[m][k] = (s_{k+1}^{m+1} - s_k^{m+1}) / (m+1), m in [0, n], k in [0, n-1)
    std::vector<double> u((n + 1) * (n - 1));
    for (int m = 0; m <= n; ++m)
        for (int k = 0; k < n - 1; ++k) {
            double a = std::exp((double)(m + 1) * std::log(s[k + 1])) / (double)(m + 1);
            double b = std::exp((double)(m + 1) * std::log(s[k])) / (double)(m + 1);
            u[m * (n - 1) + k] = a - b;
        }
    std::vector<double> W1(n * (n - 1), 0.0), W2(n * (n - 1), 0.0);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n - 1; ++k) {
            double a1 = 0, a2 = 0;
            for (int j = 0; j < n; ++j) {
                a1 += V[i * n + j] * u[(j + 1) * (n - 1) + k];
                a2 += V[i * n + j] * u[j * (n - 1) + k];
            }
            W1[i * (n - 1) + k] = a1;
            W2[i * (n - 1) + k] = a2;
        }
    std::vector<double> p(N - 1);
    {
        int idx = 0;
        for (int j = 1; j < n; ++j) p[idx++] = 1.0 / j;
        for (int j = 0; j < N - 2 * (n - 1) - 1; ++j) p[idx++] = 1.0 / (n - 1);
        for (int j = n - 1; j >= 1; --j) p[idx++] = 1.0 / j;
    }
    std::vector<double> w(N, 0.0);
    for (int j = 0; j <= N - n; ++j) {
        double a = t[j + n - 1] - t[j];
        for (int k = 0; k < n - 1; ++k) {
            double pk = p[j + k];
            for (int i = 0; i < n; ++i)
                w[j + i] += pk * (a * a * W1[i * (n - 1) + k] + a * t[j] * W2[i * (n - 1) + k]);
        }
    }
    double w40 = w[N - 40];
    for (int j = N - 40; j < N; ++j) w[j] = w40 / (double)(N - 40) * (double)j;
    return w;
}

double sinc(double x) {
    if (x == 0.0) return 1.0;
    double y = M_PI * x;
    return std::sin(y) / y;
}

}  // namespace

namespace lp {

int filter_code(const char* name) {
    if (!name) return 0;
    std::string f(name);
    for (auto& ch : f) ch = (char)std::tolower((unsigned char)ch);
    if (f == "none") return 0;
    static const char* names[] = {"ramp", "shepp-logan", "cosine", "cosine2", "hamming", "hann", "parzen"};
    for (int i = 0; i < 7; ++i)
        if (f == names[i]) return i + 1;
    return -1;
}

// take_filter (initsadj.py:146-173)
std::vector<float> take_filter(int N, int code) {
    const double d = 0.5;
    int Ne = 4 * N;
    int nt = Ne / 2 + 1;
    std::vector<double> t(nt);
    for (int j = 0; j < nt; ++j) t[j] = (double)j / (double)Ne;
    std::vector<double> w = wint(12, t);
    std::vector<double> wfa(nt);
    for (int j = 0; j < nt; ++j) {
        double base = Ne * 0.5 * w[j];
        double td = t[j] / d;
        double v = base;
        switch (code) {
            case 1: break;
            case 2: v = base * sinc(t[j] / (2 * d)) * (td <= 2 ? 1.0 : 0.0); break;
            case 3: v = base * std::cos(M_PI * t[j] / (2 * d)) * (td <= 1 ? 1.0 : 0.0); break;
            case 4: { double c = std::cos(M_PI * t[j] / (2 * d)); v = base * c * c * (td <= 1 ? 1.0 : 0.0); } break;
            case 5: v = base * (0.54 + 0.46 * std::cos(M_PI * t[j] / d)) * (td <= 1 ? 1.0 : 0.0); break;
            case 6: v = base * (1 + std::cos(M_PI * t[j] / d)) / 2.0 * (td <= 1 ? 1.0 : 0.0); break;
            case 7: { double q = 1 - td; v = base * q * q * q * (td <= 1 ? 1.0 : 0.0); } break;
        }
        wfa[j] = v * (v >= 0 ? 1.0 : 0.0);
    }
    std::vector<float> out(Ne);
    // [flipud(wfa[1:]), 2 wfa[0], wfa[1:]][:-1]
    int idx = 0;
    for (int j = nt - 1; j >= 1; --j) out[idx++] = (float)wfa[j];
    out[idx++] = (float)(2 * wfa[0]);
    for (int j = 1; j < nt - 1; ++j) out[idx++] = (float)wfa[j];
    return out;
}

}  // namespace lp

struct lp_pre {
    std::vector<float> glpars, fwdparamsf, adjparamsf;
    std::vector<int32_t> fwdparsi, adjparsi;
};

namespace {

struct Gl {
    int N0, Nproj0, Nslices, N, os, P, Ntheta, Nrho;
    double cor, beta, aR, am, g;
    bool cubic;
    std::vector<double> proj, s;
    std::vector<float> thsp, rhosp;
    std::vector<cd> B3th, B3rho;
};

void create_gl(Gl& G, int N0, int Nproj, int Nslices, double cor, bool cubic, std::vector<float>& glpars) {
    G.N0 = N0; G.Nproj0 = Nproj; G.Nslices = Nslices; G.cor = cor; G.cubic = cubic;
    G.beta = M_PI / 3;
    G.N = lp::padded_size(N0, cor);
    G.os = lp::os_angles(G.N, Nproj);
    G.P = G.os * Nproj;
    G.proj.resize(G.P);
    for (int i = 0; i < G.P; ++i) G.proj[i] = (double)i * M_PI / (double)G.P - G.beta / 2;
    G.s = linspace(-1, 1, G.N);
    double hb = G.beta / 2;
    G.aR = std::sin(hb) / (1 + std::sin(hb));
    G.am = (std::cos(hb) - std::sin(hb)) / (1 + std::sin(hb));
    {   // osg (initsgl.py:73-77)
        std::vector<double> t = linspace(-M_PI / 2, M_PI / 2, 1000);
        double best = -INFINITY;
        for (double tt : t) {
            double wr = G.aR * std::cos(tt) + (1 - G.aR), wi = G.aR * std::sin(tt);
            double v = std::log(std::hypot(wr, wi)) + std::log(std::cos(hb - std::atan2(wi, wr)));
            if (v > best) best = v;
        }
        G.g = best;
    }
    G.Ntheta = G.N;
    G.Nrho = 2 * G.N;
    double dtheta = (2 * G.beta) / G.Ntheta;
    double drho = (G.g - std::log(G.am)) / G.Nrho;
    double dth32 = (double)(float)dtheta;
    G.thsp.resize(G.Ntheta);
    for (int j = 0; j < G.Ntheta; ++j) G.thsp[j] = (float)(((double)j - G.Ntheta / 2) * dth32);
    G.rhosp.resize(G.Nrho);
    for (int j = 0; j < G.Nrho; ++j) G.rhosp[j] = (float)(((double)j - G.Nrho) * drho);
    G.B3th = spline_b3_hat(G.thsp);
    G.B3rho = spline_b3_hat(G.rhosp);
    glpars.assign(10 + (size_t)G.Nrho * G.Ntheta, 0.f);
    float head[10] = {(float)G.N, (float)N0, (float)G.Ntheta, (float)G.Nrho, 3.f, (float)G.P,
                      (float)Nslices, (float)cor, (float)G.os, cubic ? 1.f : 0.f};
    std::memcpy(glpars.data(), head, sizeof(head));
    for (int i = 0; i < G.Nrho; ++i) {
        float e = (float)std::exp((double)G.rhosp[i]);
        for (int j = 0; j < G.Ntheta; ++j) glpars[10 + (size_t)i * G.Ntheta + j] = e;
    }
}

void append_fz(std::vector<float>& dst, const Gl& G, const std::vector<cd>& half, double cst) {
    int H = G.Ntheta / 2 + 1;
    for (int i = 0; i < G.Nrho; ++i)
        for (int c = 0; c < H; ++c) {
            cd v = half[(size_t)i * H + c] * cst;
            if (G.cubic) v /= G.B3rho[i] * G.B3th[c];
            dst.push_back((float)v.real());
            dst.push_back((float)v.imag());
        }
}

void create_fwd(const Gl& G, lp_pre& out) {
    std::vector<cd> fz = fzeta_half(G.Ntheta, G.Nrho, 2 * G.beta, G.g - std::log(G.am), false);
    size_t npts = (size_t)G.Nrho * G.Ntheta;
    std::vector<double> tmp1(npts), tmp2(npts);
    for (int i = 0; i < G.Nrho; ++i) {
        double er = std::exp((double)G.rhosp[i]);
        for (int j = 0; j < G.Ntheta; ++j) {
            tmp1[(size_t)i * G.Ntheta + j] = er * std::cos((double)G.thsp[j]);
            tmp2[(size_t)i * G.Ntheta + j] = er * std::sin((double)G.thsp[j]);
        }
    }
    std::vector<int32_t> cids;
    std::vector<std::vector<double>> c1(3), c2(3);
    for (int k = 0; k < 3; ++k) {
        double phi = k * G.beta + G.beta / 2;
        double cp = std::cos(phi), sp = std::sin(phi);
        std::vector<int32_t> ck;
        for (size_t t = 0; t < npts; ++t) {
            double a = ((tmp1[t] - (1 - G.aR)) * cp - tmp2[t] * sp) / G.aR;
            double b = ((tmp1[t] - (1 - G.aR)) * sp + tmp2[t] * cp) / G.aR;
            b *= -1;
            if (a * a + b * b <= 1) {
                ck.push_back((int32_t)t);
                c1[k].push_back((a + 1) / 2 * (G.N - 1));
                c2[k].push_back((b + 1) / 2 * (G.N - 1));
            }
        }
        if (k > 0 && ck != cids) throw std::runtime_error("C->LP disk membership differs between spans");
        cids.swap(ck);
    }
    // polar grid split by spans (initsfwd.py:33-40), row-major [angle][s]
    std::vector<std::vector<int32_t>> pids(3);
    std::vector<std::vector<double>> p1(3), p2(3);
    double t0 = (double)G.thsp[0], t1 = (double)G.thsp[G.Ntheta - 1];
    double r0 = (double)G.rhosp[0], r1 = (double)G.rhosp[G.Nrho - 1];
    for (int k = 0; k < 3; ++k) {
        double lo = k * G.beta - G.beta / 2, hi = k * G.beta + G.beta / 2;
        for (int a = 0; a < G.P; ++a) {
            double th = G.proj[a];
            if (!(th >= lo && th < hi)) continue;
            double th00 = th - k * G.beta;
            for (int j = 0; j < G.N; ++j) {
                pids[k].push_back((int32_t)((size_t)a * G.N + j));
                double rr = std::log(G.s[j] * G.aR + (1 - G.aR) * std::cos(th00));
                p1[k].push_back((th00 - t0) / (t1 - t0) * (G.Ntheta - 1));
                p2[k].push_back((rr - r0) / (r1 - r0) * (G.Nrho - 1));
            }
        }
    }
    auto& pi = out.fwdparsi;
    auto& pf = out.fwdparamsf;
    for (int k = 0; k < 3; ++k) pi.push_back((int32_t)pids[k].size());
    for (int k = 0; k < 3; ++k) pi.insert(pi.end(), pids[k].begin(), pids[k].end());
    pi.push_back((int32_t)cids.size());
    pi.insert(pi.end(), cids.begin(), cids.end());
    for (auto* arr : {&c1, &c2, &p1, &p2})
        for (int k = 0; k < 3; ++k)
            for (double v : (*arr)[k]) pf.push_back((float)v);
    double cst = std::sqrt((double)G.N * G.os / G.P) * M_PI / 4 / G.aR / std::sqrt(2.0);
    append_fz(pf, G, fz, cst);
}

void create_adj(const Gl& G, int fcode, lp_pre& out) {
    std::vector<cd> fz = fzeta_half(G.Ntheta, G.Nrho, 2 * G.beta, G.g - std::log(G.am), true);
    int N = G.N;
    std::vector<double> lin = linspace(-1, 1, N);
    std::vector<int32_t> cids;
    std::vector<double> xs, ys;
    for (int r = 0; r < N; ++r)
        for (int c = 0; c < N; ++c) {
            double x1 = lin[c], x2 = -lin[r];
            if (x1 * x1 + x2 * x2 <= 1) { cids.push_back(r * N + c); xs.push_back(x1); ys.push_back(x2); }
        }
    double t0 = (double)G.thsp[0], t1 = (double)G.thsp[G.Ntheta - 1];
    double r0 = (double)G.rhosp[0], r1 = (double)G.rhosp[G.Nrho - 1];
    std::vector<std::vector<double>> C1(3), C2(3);
    for (int k = 0; k < 3; ++k) {
        double phi = k * G.beta + G.beta / 2;
        double cp = std::cos(phi), sp = std::sin(phi);
        for (size_t t = 0; t < cids.size(); ++t) {
            double z1 = G.aR * (xs[t] * cp + ys[t] * sp) + (1 - G.aR);
            double z2 = G.aR * (-xs[t] * sp + ys[t] * cp);
            double th = std::atan2(z2, z1);
            double rh = std::log(std::sqrt(z1 * z1 + z2 * z2));
            C1[k].push_back((th - t0) / (t1 - t0) * (G.Ntheta - 1));
            C2[k].push_back((rh - r0) / (r1 - r0) * (G.Nrho - 1));
        }
    }
    // log-polar -> polar (initsadj.py:40-72)
    float hb32n = (float)(-G.beta / 2), hb32p = (float)(G.beta / 2);
    std::vector<int32_t> lpids, wr, wl;
    std::vector<double> lz2n, zr, zl;
    double drho = (double)G.rhosp[1] - (double)G.rhosp[0];
    double lam = std::log(G.am);
    for (int i = 0; i < G.Nrho; ++i) {
        double z2 = std::exp((double)G.rhosp[i]);
        double lz = std::log(z2);
        for (int j = 0; j < G.Ntheta; ++j) {
            float z1f = G.thsp[j];
            double z1 = (double)z1f;
            bool inspan = (z1f >= hb32n) && (z1f < hb32p);
            int32_t id = i * G.Ntheta + j;
            double z2n = (z2 - (1 - G.aR) * std::cos(z1)) / G.aR;
            if (inspan && std::fabs(z2n) <= 1) { lpids.push_back(id); lz2n.push_back(z2n); }
            if (lz > G.g) {
                double v = (std::exp(lz + lam - G.g) - (1 - G.aR) * std::cos(z1)) / G.aR;
                if (inspan && std::fabs(v) <= 1) { wr.push_back(id); zr.push_back(v); }
            }
            if (lz < lam - G.g + drho) {
                double v = (std::exp(lz - lam + G.g) - (1 - G.aR) * std::cos(z1)) / G.aR;
                if (inspan && std::fabs(v) <= 1) { wl.push_back(id); zl.push_back(v); }
            }
        }
    }
    std::vector<int32_t> wids(wr);
    wids.insert(wids.end(), wl.begin(), wl.end());
    std::vector<double> wz2n(zr);
    wz2n.insert(wz2n.end(), zl.begin(), zl.end());
    // angle index map (initsadj.py:74-95)
    int pfirst[3], plast[3], plen[3];
    for (int k = 0; k < 3; ++k) {
        double lo = k * G.beta - G.beta / 2, hi = k * G.beta + G.beta / 2;
        pfirst[k] = -1; plast[k] = -1; plen[k] = 0;
        for (int a = 0; a < G.P; ++a)
            if (G.proj[a] >= lo && G.proj[a] < hi) {
                if (pfirst[k] < 0) pfirst[k] = a;
                plast[k] = a;
                plen[k]++;
            }
        if (plen[k] < 2) throw std::runtime_error("a span holds fewer than two projection angles");
    }
    double proj0[3], projl[3];
    for (int k = 0; k < 3; ++k) { proj0[k] = G.proj[pfirst[k]]; projl[k] = G.proj[plast[k]] - G.proj[pfirst[k]]; }
    double projp = (G.P - 1) / (proj0[2] + projl[2] - proj0[0]);
    auto theta_of = [&](int32_t id) { return (double)G.thsp[id % G.Ntheta]; };
    std::vector<std::vector<double>> L1(3), L2(3), W1(3), W2(3);
    for (int k = 0; k < 3; ++k) {
        for (size_t t = 0; t < lpids.size(); ++t) {
            double a = theta_of(lpids[t]) + k * G.beta;
            L1[k].push_back((a - proj0[k]) / projl[k] * (plen[k] - 1) + (proj0[k] - proj0[0]) * projp);
            L2[k].push_back((lz2n[t] + 1) / 2 * (N - 1));
        }
        for (size_t t = 0; t < wids.size(); ++t) {
            double a = theta_of(wids[t]) + k * G.beta;
            W1[k].push_back((a - proj0[k]) / projl[k] * (plen[k] - 1) + (proj0[k] - proj0[0]) * projp);
            W2[k].push_back((wz2n[t] + 1) / 2 * (N - 1));
        }
    }
    auto& pi = out.adjparsi;
    auto& pf = out.adjparamsf;
    pi.push_back((int32_t)cids.size());
    pi.push_back((int32_t)lpids.size());
    pi.push_back((int32_t)wids.size());
    pi.insert(pi.end(), cids.begin(), cids.end());
    pi.insert(pi.end(), lpids.begin(), lpids.end());
    pi.insert(pi.end(), wids.begin(), wids.end());
    for (auto* arr : {&C1, &C2, &L1, &L2, &W1, &W2})
        for (int k = 0; k < 3; ++k)
            for (double v : (*arr)[k]) pf.push_back((float)v);
    double cst = (double)(N + 1) * (N - 1) / ((double)N * N) / 2 * std::sqrt((double)G.os * G.P / N / 2);
    append_fz(pf, G, fz, cst);
    if (fcode > 0) {
        std::vector<float> filt = lp::take_filter(N, fcode);
        pf.insert(pf.end(), filt.begin(), filt.end());
        pi.push_back(1);
    } else {
        pi.push_back(0);
    }
}

}  // namespace

namespace lp {

int padded_size(int N0, double cor) {
    return (int)(std::ceil((N0 + std::fabs(N0 / 2.0 - cor) * 2.0) / 16.0) * 16);
}

int os_angles(int N, int Nproj) {
    double v = std::nearbyint(3.0 * N / 2.0 / Nproj);  // round-half-to-even like Python round()
    return (int)std::max(v, 1.0);
}

}  // namespace lp

extern "C" {

const char* lp_pre_error(void) { return g_pre_err.c_str(); }

int lp_pre_create(int N0, int Nproj, int Nslices, double cor, int cubic, const char* filter, int want_fwd,
                  lp_pre** out) {
    g_pre_err.clear();
    if (!out) { g_pre_err = "null output pointer"; return LP_ERR_ARG; }
    *out = nullptr;
    if (N0 < 16 || Nproj < 3 || Nslices < 1) { g_pre_err = "need N0 >= 16, Nproj >= 3, Nslices >= 1"; return LP_ERR_ARG; }
    if (!(cor > -1e6 && cor < 1e6)) { g_pre_err = "cor out of range"; return LP_ERR_ARG; }
    int fcode = lp::filter_code(filter);
    if (fcode < 0) { g_pre_err = std::string("unknown filter type '") + (filter ? filter : "") + "'"; return LP_ERR_ARG; }
    try {
        std::unique_ptr<lp_pre> p(new lp_pre());
        Gl G;
        create_gl(G, N0, Nproj, Nslices, cor, cubic != 0, p->glpars);
        if (G.N / 2 - (int)cor < 0 || G.N / 2 + N0 - (int)cor > G.N) { g_pre_err = "cor puts the detector outside the padded grid"; return LP_ERR_ARG; }
        if (want_fwd) create_fwd(G, *p);
        create_adj(G, fcode, *p);
        *out = p.release();
        return LP_OK;
    } catch (const std::bad_alloc&) {
        g_pre_err = "out of host memory in precompute";
        return LP_ERR_ALLOC;
    } catch (const std::exception& e) {
        g_pre_err = e.what();
        return LP_ERR_ARG;
    }
}

int lp_pre_sizes(const lp_pre* p, int64_t* sizes) {
    if (!p || !sizes) return LP_ERR_ARG;
    sizes[0] = (int64_t)p->glpars.size();
    sizes[1] = (int64_t)p->fwdparsi.size();
    sizes[2] = (int64_t)p->fwdparamsf.size();
    sizes[3] = (int64_t)p->adjparsi.size();
    sizes[4] = (int64_t)p->adjparamsf.size();
    return LP_OK;
}

int lp_pre_copy(const lp_pre* p, float* glpars, int32_t* fwdparsi, float* fwdparamsf, int32_t* adjparsi,
                float* adjparamsf) {
    if (!p) return LP_ERR_ARG;
    if (glpars) std::memcpy(glpars, p->glpars.data(), p->glpars.size() * 4);
    if (fwdparsi) std::memcpy(fwdparsi, p->fwdparsi.data(), p->fwdparsi.size() * 4);
    if (fwdparamsf) std::memcpy(fwdparamsf, p->fwdparamsf.data(), p->fwdparamsf.size() * 4);
    if (adjparsi) std::memcpy(adjparsi, p->adjparsi.data(), p->adjparsi.size() * 4);
    if (adjparamsf) std::memcpy(adjparamsf, p->adjparamsf.data(), p->adjparamsf.size() * 4);
    return LP_OK;
}

void lp_pre_destroy(lp_pre* p) { delete p; }

}  // extern "C"
