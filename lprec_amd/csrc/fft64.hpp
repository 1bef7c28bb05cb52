// Host FP64 FFT used by the precompute (kernel spectra, filter kernels).
// Mixed-radix Stockham for lengths whose prime factors are <= 64, Bluestein
// (chirp-z over a power-of-two Stockham) for everything else.
#pragma once
#include <cmath>
#include <complex>
#include <vector>

namespace lpfft {

using cd = std::complex<double>;

inline std::vector<int> factorize(int n) {
    std::vector<int> f;
    while (n % 4 == 0) { f.push_back(4); n /= 4; }
    while (n % 2 == 0) { f.push_back(2); n /= 2; }
    for (int p = 3; p * p <= n; p += 2)
        while (n % p == 0) { f.push_back(p); n /= p; }
    if (n > 1) f.push_back(n);
    return f;
}

// Unnormalised DFT, sign = -1 forward, +1 inverse.
class Plan {
public:
    explicit Plan(int n, int sign = -1) : n_(n), sign_(sign) {
        fac_ = factorize(n);
        bool small = true;
        for (int r : fac_) if (r > 64) small = false;
        if (n <= 1) { bluestein_ = false; return; }
        if (small) {
            bluestein_ = false;
            tw_.resize(n);
            for (int j = 0; j < n; ++j) {
                double a = sign_ * 2.0 * M_PI * (double)j / (double)n;
                tw_[j] = cd(std::cos(a), std::sin(a));
            }
        } else {
            bluestein_ = true;
            m_ = 1;
            while (m_ < 2 * n - 1) m_ <<= 1;
            chirp_.resize(n);
            for (long long j = 0; j < n; ++j) {
                long long jj = (j * j) % (2LL * n);
                double a = sign_ * M_PI * (double)jj / (double)n;
                chirp_[j] = cd(std::cos(a), std::sin(a));
            }
            fwd_.reset(new Plan(m_, -1));
            inv_.reset(new Plan(m_, +1));
            std::vector<cd> b(m_, cd(0, 0));
            b[0] = std::conj(chirp_[0]);
            for (int j = 1; j < n; ++j) { b[j] = std::conj(chirp_[j]); b[m_ - j] = std::conj(chirp_[j]); }
            fwd_->exec(b);
            bhat_ = b;
        }
    }
    Plan(const Plan&) = delete;
    Plan& operator=(const Plan&) = delete;

    void exec(std::vector<cd>& x) const {
        if (n_ <= 1) return;
        if (!bluestein_) { stockham(x); return; }
        std::vector<cd> a(m_, cd(0, 0));
        for (int j = 0; j < n_; ++j) a[j] = x[j] * chirp_[j];
        fwd_->exec(a);
        for (int j = 0; j < m_; ++j) a[j] *= bhat_[j];
        inv_->exec(a);
        double s = 1.0 / (double)m_;
        for (int j = 0; j < n_; ++j) x[j] = a[j] * s * chirp_[j];
    }
    int size() const { return n_; }

private:
    void stockham(std::vector<cd>& x) const {
        std::vector<cd> y(n_);
        cd* in = x.data();
        cd* out = y.data();
        int ns = 1;
        std::vector<cd> v(64), w(64);
        for (int r : fac_) {
            int nb = n_ / r;
            int step = n_ / r;
            for (int i = 0; i < nb; ++i) {
                int k = i % ns;
                int tstep = (int)((long long)k * n_ / ((long long)ns * r));
                for (int q = 0; q < r; ++q) {
                    cd t = in[i + q * step];
                    if (q) t *= tw_[(long long)q * tstep % n_];
                    v[q] = t;
                }
                for (int p = 0; p < r; ++p) {
                    cd acc(0, 0);
                    for (int q = 0; q < r; ++q) acc += v[q] * tw_[(long long)p * q * (n_ / r) % n_];
                    w[p] = acc;
                }
                int j = (i / ns) * ns * r + k;
                for (int p = 0; p < r; ++p) out[j + p * ns] = w[p];
            }
            ns *= r;
            std::swap(in, out);
        }
        if (in != x.data()) for (int j = 0; j < n_; ++j) x[j] = in[j];
    }

    int n_, sign_;
    std::vector<int> fac_;
    std::vector<cd> tw_;
    bool bluestein_ = false;
    int m_ = 0;
    std::vector<cd> chirp_, bhat_;
    std::unique_ptr<Plan> fwd_, inv_;
};

}  // namespace lpfft
