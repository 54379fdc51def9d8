// lagrange.cpp -- bootstrapping keys in the upstream FFT-domain (LagrangeHalfCPolynomial) order.
//
// Upstream tfhe, and the reference's copies of its conversions, hold a torus polynomial t of
// degree < N by its values at the upper-half roots of X^N + 1, in torus units:
//     L_i = sum_j (t_j / 2^32) w^{(2i+1) j},   w = e^{i pi / N},   i < N/2.
// Int → Lagrange: tGswFFTExternMulToTLwe.cpp:97-130 (2N-point DFT of [a/2, -a/2], conjugated,
// odd outputs).  Lagrange → torus: TorusPolynomial_fft.cpp:4-43 (the 2N-point embedding of the
// values and their conjugates, scaled by 2^32/N), which fixes the torus-unit scale of the key.
//
// This engine's device key is built from the time-domain key (exact integers), so an FFT-domain
// key is brought back to the time domain first.  With all N roots, L_{N-1-i} = conj(L_i), so
//     t_j / 2^32 = (1/N) sum_{i<N} L_i w^{-(2i+1) j} = (2/N) Re( w^{-j} sum_{i<N/2} L_i e^{-2 pi i ij/N} ),
// one N-point transform of the zero-padded values per polynomial, rounded to the nearest integer.
// A value further than 1/4 from an integer means the input is not the transform of a torus32
// polynomial; the conversion then fails rather than guess.
#include <math.h>
#include <string.h>

#include <complex>
#include <string>
#include <vector>

#include "internal.h"

namespace tfhe {
namespace {

using cd = std::complex<double>;

// In-place radix-2 DFT of a power-of-two size n: X_k = sum_j x_j e^{sgn 2 pi i jk / n}.
class HostDft {
public:
    explicit HostDft(int n) : n_(n), tw_(n / 2), rev_(n) {
        const long double pi = 3.141592653589793238462643383279502884L;
        for (int k = 0; k < n / 2; ++k) {
            const long double a = 2.0L * pi * k / n;
            tw_[k] = cd((double)cosl(a), (double)sinl(a));
        }
        int lg = 0;
        while ((1 << lg) < n) ++lg;
        for (int i = 0; i < n; ++i) {
            int r = 0;
            for (int b = 0; b < lg; ++b) r |= ((i >> b) & 1) << (lg - 1 - b);
            rev_[i] = r;
        }
    }
    void run(cd* x, int sgn) const {
        for (int i = 0; i < n_; ++i)
            if (i < rev_[i]) std::swap(x[i], x[rev_[i]]);
        for (int len = 2; len <= n_; len <<= 1) {
            const int step = n_ / len;
            for (int s = 0; s < n_; s += len)
                for (int k = 0; k < len / 2; ++k) {
                    cd w = tw_[k * step];
                    if (sgn < 0) w = std::conj(w);
                    const cd u = x[s + k], v = x[s + k + len / 2] * w;
                    x[s + k] = u + v;
                    x[s + k + len / 2] = u - v;
                }
        }
    }

private:
    int n_;
    std::vector<cd> tw_;
    std::vector<int> rev_;
};

bool flat_ok(const tfhe_flat_params* p) {
    if (!p || p->n <= 0 || p->k < 1 || p->l < 1 || p->N < 4 || p->N > 65536) return false;
    return (p->N & (p->N - 1)) == 0;
}

size_t bk_polys(const tfhe_flat_params* p) { return (size_t)p->n * (p->k + 1) * p->l * (p->k + 1); }

// torus polynomial from its upstream FFT-domain values (N/2 complex, re/im interleaved)
bool torus_from_values(const HostDft& dft, int N, const double* lag, int32_t* t, std::vector<cd>& buf,
                       double* worst) {
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int i = 0; i < N / 2; ++i) buf[i] = cd(lag[2 * i], lag[2 * i + 1]);
    for (int i = N / 2; i < N; ++i) buf[i] = cd(0.0, 0.0);
    dft.run(buf.data(), -1);
    const double scale = 2.0 / N * 4294967296.0;
    for (int j = 0; j < N; ++j) {
        const long double a = -pi * j / N;
        const cd w((double)cosl(a), (double)sinl(a));
        const double x = (w * buf[j]).real() * scale;
        if (!isfinite(x) || fabs(x) > 4.0 * 4294967296.0) return false;
        const double r = nearbyint(x);
        const double d = fabs(x - r);
        if (d > *worst) *worst = d;
        if (d > 0.25) return false;
        t[j] = (int32_t)(uint32_t)(int64_t)r;   // wraps mod 2^32
    }
    return true;
}

void values_from_torus(const HostDft& dft, int N, const int32_t* t, double* lag, std::vector<cd>& buf) {
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int j = 0; j < N; ++j) {
        const long double a = pi * j / N;
        buf[j] = cd((double)cosl(a), (double)sinl(a)) * ((double)t[j] / 4294967296.0);
    }
    dft.run(buf.data(), +1);
    for (int i = 0; i < N / 2; ++i) {
        lag[2 * i] = buf[i].real();
        lag[2 * i + 1] = buf[i].imag();
    }
}

int from_lagrange(const tfhe_flat_params* p, const double* lagrange, int32_t* bk) {
    const int N = p->N;
    HostDft dft(N);
    std::vector<cd> buf(N);
    double worst = 0.0;
    const size_t np = bk_polys(p);
    for (size_t q = 0; q < np; ++q)
        if (!torus_from_values(dft, N, lagrange + q * N, bk + q * N, buf, &worst)) {
            set_error("polynomial " + std::to_string(q) +
                      " of the FFT-domain key is not the transform of a torus32 polynomial");
            return TFHE_E_ARG;
        }
    return TFHE_OK;
}

}  // namespace
}  // namespace tfhe

using namespace tfhe;

extern "C" int tfhe_bk_from_lagrange(const tfhe_flat_params* p, const double* lagrange, int32_t* bk) {
    if (!flat_ok(p) || !lagrange || !bk) return lib_arg_fail("tfhe_bk_from_lagrange: bad parameters or null array");
    return from_lagrange(p, lagrange, bk);
}

extern "C" int tfhe_bk_to_lagrange(const tfhe_flat_params* p, const int32_t* bk, double* lagrange) {
    if (!flat_ok(p) || !lagrange || !bk) return lib_arg_fail("tfhe_bk_to_lagrange: bad parameters or null array");
    const int N = p->N;
    HostDft dft(N);
    std::vector<cd> buf(N);
    const size_t np = bk_polys(p);
    for (size_t q = 0; q < np; ++q) values_from_torus(dft, N, bk + q * N, lagrange + q * N, buf);
    return TFHE_OK;
}

extern "C" int tfhe_gpu_key_create_lagrange(const tfhe_flat_params* p, const double* lagrange, const int32_t* ksk,
                                            int32_t prec_mask, tfhe_gpu_key** out) {
    if (!flat_ok(p) || !lagrange || !ksk || !out)
        return lib_arg_fail("tfhe_gpu_key_create_lagrange: bad parameters or null argument");
    std::vector<int32_t> bk(bk_polys(p) * p->N);
    const int rc = from_lagrange(p, lagrange, bk.data());
    if (rc) return rc;
    return tfhe_gpu_key_create(p, bk.data(), ksk, prec_mask, out);
}

namespace tfhe {

// The bk of an upstream-shaped LweBootstrappingKeyFFT whose polynomials carry their values
// (LagrangeHalfCPolynomial.data = N/2 complex doubles), as flattened at host_to_fpga.h:65-72.
int bk_from_lagrange_struct(const LweBootstrappingKeyFFT* bkfft, const tfhe_flat_params* p, int32_t* bk) {
    const int kpl = (p->k + 1) * p->l, M = p->N / 2;
    std::vector<double> lag(bk_polys(p) * p->N);
    size_t o = 0;
    for (int i = 0; i < p->n; ++i)
        for (int r = 0; r < kpl; ++r)
            for (int q = 0; q <= p->k; ++q) {
                const void* d = bkfft->bkFFT[i].all_samples[r].a[q].data;
                if (!d) {
                    set_error("FFT-domain key polynomial without values (LagrangeHalfCPolynomial.data is NULL)");
                    return TFHE_E_ARG;
                }
                memcpy(lag.data() + o, d, sizeof(double) * 2 * M);
                o += (size_t)2 * M;
            }
    return from_lagrange(p, lag.data(), bk);
}

}  // namespace tfhe
