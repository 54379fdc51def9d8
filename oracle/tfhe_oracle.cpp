/*
 * tfhe_oracle.cpp -- CPU oracle for the TFHE gate-bootstrapping path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library.
 *
 * Reference files restated here (paths relative to the reference root):
 *   kernel/src/numeric-functions.cpp                 torus numerics, global RNG
 *   kernel/src/lwe-functions.cpp                     LWE keygen/encrypt/phase/arith
 *   kernel/src/lwe-keyswitch-functions.cpp           KSK generation and keyswitch
 *   kernel/src/tfhe_bootstrap_woKS_FFT/tfhe_bootstrap_woKS_FFT.cpp   modswitch, test vector
 *   kernel/src/tfhe_blindRotateAndExtract/mulByXaiOrCopy.cpp         X^a rotation
 *   kernel/src/tfhe_blindRotate_FFT/tfhe_MuxRotate_FFT.cpp           (X^a-1), MuxRotate, AddTo
 *   kernel/src/tfhe_blindRotate_FFT/tGswTorus32PolynomialDecompH.cpp gadget decomposition
 *   kernel/src/tfhe_blindRotate_FFT/tGswFFTExternMulToTLwe.cpp       external product
 *   kernel/src/tfhe_blindRotate_FFT/TorusPolynomial_fft.cpp          FFT -> torus conversion
 *   kernel/src/tfhe_blindRotate_FFT/tfhe_blindRotate_FFT.cpp         blind-rotate loop
 *   kernel/src/tfhe_blindRotate_FFT/tfhe_blindRotate_FFT_tb.cpp      gate pre-combination, MU
 * TLWE/TGSW key generation, sample extraction and the gate constants other
 * than OR live only in the upstream `tfhe` library, which the reference links
 * but does not vendor; they are restated from the CGGI construction that the
 * reference's call sites use (test_bench/server/cloud.c, tb :267-325).
 */
#include "tfhe_oracle.h"

#include <chrono>
#include <climits>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <mutex>
#include <random>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef std::complex<double> cplx;

static int g_fault = 0;
extern "C" void orc_set_fault(int32_t fault) { g_fault = fault; }

/* ------------------------------------------------------------------------ */
/* torus numerics: kernel/src/numeric-functions.cpp                         */
/* ------------------------------------------------------------------------ */

/* numeric-functions.cpp:11-13 -- one global engine and a global uniform torus distribution */
static std::default_random_engine g_gen;
static std::uniform_int_distribution<int32_t> g_unif_torus(INT32_MIN, INT32_MAX);

/* numeric-functions.cpp:16-19 */
extern "C" void orc_set_seed(const uint32_t* values, int32_t size) {
    std::seed_seq s(values, values + size);
    g_gen.seed(s);
}

/* numeric-functions.cpp:32-34: fractional part of d scaled by 2^32, via int64 truncation */
extern "C" int32_t orc_dtot32(double d) {
    const double two32 = 4294967296.0;
    return (int32_t)(int64_t)((d - (double)(int64_t)d) * two32);
}

/* numeric-functions.cpp:36-38 */
extern "C" double orc_t32tod(int32_t x) { return (double)x / 4294967296.0; }

/* numeric-functions.cpp:22-27: a fresh normal_distribution object per draw */
extern "C" int32_t orc_gaussian32(int32_t message, double sigma) {
    std::normal_distribution<double> dist(0., sigma);
    double err = dist(g_gen);
    return (int32_t)((uint32_t)message + (uint32_t)orc_dtot32(err));
}

extern "C" int32_t orc_uniform_torus32(void) { return g_unif_torus(g_gen); }

static int32_t uniform_bit() {
    std::uniform_int_distribution<int32_t> d(0, 1);  /* lwe-functions.cpp:23 */
    return d(g_gen);
}

/* numeric-functions.cpp:45-53 */
extern "C" int32_t orc_approx_phase(int32_t phase, int32_t Msize) {
    uint64_t interv = ((UINT64_C(1) << 63) / (uint64_t)Msize) * 2;
    uint64_t half = interv / 2;
    uint64_t p64 = ((uint64_t)(uint32_t)phase << 32) + half;
    p64 -= p64 % interv;
    return (int32_t)(uint32_t)(p64 >> 32);
}

/* numeric-functions.cpp:59-65 (same as tfhe_bootstrap_woKS_FFT.cpp:3-10) */
extern "C" int32_t orc_modswitch_from_torus32(int32_t phase, int32_t Msize) {
    uint64_t interv = ((UINT64_C(1) << 63) / (uint64_t)Msize) * 2;
    uint64_t half = interv / 2;
    uint64_t p64 = ((uint64_t)(uint32_t)phase << 32) + half;
    return (int32_t)(p64 / interv);
}

/* numeric-functions.cpp:71-76 */
extern "C" int32_t orc_modswitch_to_torus32(int32_t mu, int32_t Msize) {
    uint64_t interv = ((UINT64_C(1) << 63) / (uint64_t)Msize) * 2;
    uint64_t p64 = (uint64_t)(int64_t)mu * interv;
    return (int32_t)(uint32_t)(p64 >> 32);
}

/* decomposition offset: sum over levels of halfBg * 2^(32-(p+1)Bgbit).
 * constants.h:24 pins 2164391936 for l=3, Bgbit=7. */
extern "C" uint32_t orc_decomp_offset(int32_t l, int32_t Bgbit) {
    uint32_t halfBg = 1u << (Bgbit - 1);
    uint32_t off = 0;
    for (int p = 0; p < l; ++p) off += halfBg * (1u << (32 - (p + 1) * Bgbit));
    return off;
}

/* ------------------------------------------------------------------------ */
/* LWE (kernel/src/lwe-functions.cpp)                                       */
/* ------------------------------------------------------------------------ */

/* lwe-functions.cpp:36-47: b drawn first (gaussian), then each a_i, b += a_i*s_i */
extern "C" void orc_lwe_encrypt(int32_t* out, int32_t message, double alpha, const int32_t* key, int32_t n) {
    uint32_t b = (uint32_t)orc_gaussian32(message, alpha);
    for (int i = 0; i < n; ++i) {
        out[i] = g_unif_torus(g_gen);
        b += (uint32_t)out[i] * (uint32_t)key[i];
    }
    out[n] = (int32_t)b;
}

/* lwe-functions.cpp:53-64 */
static void lwe_encrypt_external_noise(int32_t* out, int32_t message, double noise, const int32_t* key, int32_t n) {
    uint32_t b = (uint32_t)message + (uint32_t)orc_dtot32(noise);
    for (int i = 0; i < n; ++i) {
        out[i] = g_unif_torus(g_gen);
        b += (uint32_t)out[i] * (uint32_t)key[i];
    }
    out[n] = (int32_t)b;
}

/* lwe-functions.cpp:72-81 */
extern "C" int32_t orc_lwe_phase(const int32_t* sample, const int32_t* key, int32_t n) {
    uint32_t axs = 0;
    for (int i = 0; i < n; ++i) axs += (uint32_t)sample[i] * (uint32_t)key[i];
    return (int32_t)((uint32_t)sample[n] - axs);
}

/* upstream bootsSymEncrypt: message ? +1/8 : -1/8 with the LWE noise (alice.c:36 call site) */
extern "C" void orc_boots_encrypt(const orc_params* p, int32_t* out, int32_t bit, const int32_t* lwe_key) {
    int32_t mu = orc_modswitch_to_torus32(1, 8);
    orc_lwe_encrypt(out, bit ? mu : -mu, p->lwe_alpha, lwe_key, p->n);
}

/* upstream bootsSymDecrypt: phase > 0 (verify.c:27 call site) */
extern "C" int32_t orc_boots_decrypt(const orc_params* p, const int32_t* sample, const int32_t* lwe_key) {
    return orc_lwe_phase(sample, lwe_key, p->n) > 0 ? 1 : 0;
}

/* ------------------------------------------------------------------------ */
/* polynomial primitives                                                    */
/* ------------------------------------------------------------------------ */

/* mulByXaiOrCopy.cpp:25-41: res = X^a * src for a in [0, 2N) (negacyclic) */
extern "C" void orc_poly_mul_xai(int32_t* res, int32_t a, const int32_t* src, int32_t N) {
    if (a < N) {
        for (int i = 0; i < a; ++i) res[i] = -src[i - a + N];
        for (int i = a; i < N; ++i) res[i] = src[i - a];
    } else {
        int aa = a - N;
        for (int i = 0; i < aa; ++i) res[i] = src[i - aa + N];
        for (int i = aa; i < N; ++i) res[i] = -src[i - aa];
    }
}

/* tfhe_MuxRotate_FFT.cpp:5-24: res = (X^a - 1) * src */
extern "C" void orc_poly_mul_xai_minus_one(int32_t* res, int32_t a, const int32_t* src, int32_t N) {
    auto neg = [](int32_t x) { return (int32_t)(0u - (uint32_t)x); };
    auto sub = [](int32_t x, int32_t y) { return (int32_t)((uint32_t)x - (uint32_t)y); };
    const bool flip = (g_fault == 1);
    if (a < N) {
        for (int i = 0; i < a; ++i) res[i] = sub(flip ? src[i - a + N] : neg(src[i - a + N]), src[i]);
        for (int i = a; i < N; ++i) res[i] = sub(src[i - a], src[i]);
    } else {
        int aa = a - N;
        for (int i = 0; i < aa; ++i) res[i] = sub(flip ? neg(src[i - aa + N]) : src[i - aa + N], src[i]);
        for (int i = aa; i < N; ++i) res[i] = sub(neg(src[i - aa]), src[i]);
    }
}

/* tGswTorus32PolynomialDecompH.cpp:29-44: u = x + offset (uint32), d_p = ((u >> (32-(p+1)Bgbit)) & mask) - halfBg */
extern "C" void orc_decompose(int32_t* res, const int32_t* src, int32_t N, int32_t l, int32_t Bgbit) {
    const uint32_t off = (g_fault == 2) ? 0u : orc_decomp_offset(l, Bgbit);
    const uint32_t mask = (1u << Bgbit) - 1;
    const int32_t halfBg = 1 << (Bgbit - 1);
    for (int p = 0; p < l; ++p) {
        const int decal = 32 - (p + 1) * Bgbit;
        for (int j = 0; j < N; ++j) {
            uint32_t u = (uint32_t)src[j] + off;
            res[p * N + j] = (int32_t)((u >> decal) & mask) - halfBg;
        }
    }
}

/* exact negacyclic product in int64 (no reduction mod 2^32): res = ipoly * tpoly mod X^N+1.
 * |sum| <= N * |ipoly|max * 2^31, which fits int64 for the decomposition digits. */
extern "C" void orc_negacyclic_mul_i64(int64_t* res, const int32_t* ip, const int32_t* tp, int32_t N) {
    for (int j = 0; j < N; ++j) res[j] = 0;
    for (int i = 0; i < N; ++i) {
        const int64_t a = ip[i];
        if (a == 0) continue;
        for (int j = 0; j < N - i; ++j) res[i + j] += a * (int64_t)tp[j];
        for (int j = N - i; j < N; ++j) res[i + j - N] -= a * (int64_t)tp[j];
    }
}

/* torus product with uint32 wrap (upstream torusPolynomialAddMulR semantics, used by TLWE encryption) */
extern "C" void orc_negacyclic_mul_naive(int32_t* res, const int32_t* ip, const int32_t* tp, int32_t N) {
    std::vector<uint32_t> r(N, 0);
    for (int i = 0; i < N; ++i) {
        const uint32_t a = (uint32_t)ip[i];
        if (a == 0) continue;
        for (int j = 0; j < N - i; ++j) r[i + j] += a * (uint32_t)tp[j];
        for (int j = N - i; j < N; ++j) r[i + j - N] -= a * (uint32_t)tp[j];
    }
    for (int j = 0; j < N; ++j) res[j] = (int32_t)r[j];
}

/* ------------------------------------------------------------------------ */
/* keys: lwe-functions.cpp:21-27, lwe-keyswitch-functions.cpp:163-215,      */
/* upstream TLWE/TGSW keygen (CGGI)                                         */
/* ------------------------------------------------------------------------ */

/* upstream tLweSymEncryptZero + tGswAddMuIntH, applied row by row.
 * tgsw layout: [(k+1)*l][(k+1)][N]. */
static void tgsw_encrypt_int(const orc_params* p, int32_t* tgsw, int32_t message, const int32_t* tlwe_key) {
    const int N = p->N, k = p->k, l = p->l;
    const int kpl = (k + 1) * l;
    std::vector<int32_t> tmp(N);
    for (int r = 0; r < kpl; ++r) {
        int32_t* row = tgsw + (size_t)r * (k + 1) * N;
        int32_t* b = row + (size_t)k * N;
        for (int j = 0; j < N; ++j) b[j] = orc_gaussian32(0, p->bk_alpha);
        for (int q = 0; q < k; ++q) {
            int32_t* a = row + (size_t)q * N;
            for (int j = 0; j < N; ++j) a[j] = g_unif_torus(g_gen);
            orc_negacyclic_mul_naive(tmp.data(), tlwe_key + (size_t)q * N, a, N);
            for (int j = 0; j < N; ++j) b[j] = (int32_t)((uint32_t)b[j] + (uint32_t)tmp[j]);
        }
    }
    for (int bloc = 0; bloc <= k; ++bloc)
        for (int i = 0; i < l; ++i) {
            uint32_t h = 1u << (32 - (i + 1) * p->Bgbit);
            int32_t* c0 = tgsw + ((size_t)(bloc * l + i) * (k + 1) + bloc) * N;
            c0[0] = (int32_t)((uint32_t)c0[0] + (uint32_t)message * h);
        }
}

/* Draw order follows upstream new_random_gate_bootstrapping_secret_keyset:
 * LWE key, TLWE key, keyswitch key (extracted key -> LWE key), then n TGSW rows sets. */
extern "C" void orc_keygen(const orc_params* p, const uint32_t* seed, int32_t seedlen,
                           int32_t* lwe_key, int32_t* tlwe_key, int32_t* ksk, int32_t* bk) {
    orc_set_seed(seed, seedlen);
    const int n = p->n, N = p->N, k = p->k, t = p->ks_t, basebit = p->ks_basebit;
    const int base = 1 << basebit;
    for (int i = 0; i < n; ++i) lwe_key[i] = uniform_bit();          /* lwe-functions.cpp:21-27 */
    for (int i = 0; i < k * N; ++i) tlwe_key[i] = uniform_bit();     /* extracted key = TLWE key coefs */

    /* lwe-keyswitch-functions.cpp:163-215 */
    const int kN = k * N;
    const int sizeks = kN * t * (base - 1);
    std::vector<double> noise(sizeks);
    double err = 0;
    for (int i = 0; i < sizeks; ++i) {
        std::normal_distribution<double> d(0., p->lwe_alpha);
        noise[i] = d(g_gen);
        err += noise[i];
    }
    err /= sizeks;
    for (int i = 0; i < sizeks; ++i) noise[i] -= err;
    int index = 0;
    const size_t ks_stride = (size_t)(n + 1);
    for (int i = 0; i < kN; ++i)
        for (int j = 0; j < t; ++j) {
            int32_t* s0 = ksk + (((size_t)i * t + j) * base) * ks_stride;
            for (int c = 0; c <= n; ++c) s0[c] = 0;                    /* h = 0: trivial zero, :192 */
            for (int h = 1; h < base; ++h) {
                uint32_t mess = (uint32_t)tlwe_key[i] * (uint32_t)h * (1u << (32 - (j + 1) * basebit));
                lwe_encrypt_external_noise(s0 + (size_t)h * ks_stride, (int32_t)mess, noise[index], lwe_key, n);
                index += 1;
            }
        }

    /* bootstrapping key: bk[i] = TGSW(s_i) under the TLWE key */
    const size_t tgsw_sz = (size_t)(k + 1) * p->l * (k + 1) * N;
    for (int i = 0; i < n; ++i) tgsw_encrypt_int(p, bk + (size_t)i * tgsw_sz, lwe_key[i], tlwe_key);
}

/* phase polynomials of a TLWE sample: b - sum_q a_q * s_q */
extern "C" void orc_tlwe_phase(const orc_params* p, int32_t* phase, const int32_t* sample, const int32_t* tlwe_key) {
    const int N = p->N, k = p->k;
    std::vector<int32_t> tmp(N);
    for (int j = 0; j < N; ++j) phase[j] = sample[(size_t)k * N + j];
    for (int q = 0; q < k; ++q) {
        orc_negacyclic_mul_naive(tmp.data(), tlwe_key + (size_t)q * N, sample + (size_t)q * N, N);
        for (int j = 0; j < N; ++j) phase[j] = (int32_t)((uint32_t)phase[j] - (uint32_t)tmp[j]);
    }
}

/* ------------------------------------------------------------------------ */
/* FFT: folded negacyclic transform (FP64).                                  */
/* A real polynomial p of degree < N is evaluated at w^(4k+1), w = e^{i pi/N}, */
/* k < N/2, by z_j = (p_j + i p_{j+N/2}) w^j and F_k = sum_j z_j e^{+2 pi i jk/(N/2)}. */
/* These are the odd powers of w used by tGswFFTExternMulToTLwe.cpp:97-130  */
/* (the reference's LagrangeHalfC set w^(2i+1), i<N/2), up to an exact      */
/* permutation and conjugation; orc_lagrange_half_reference() maps between. */
/* ------------------------------------------------------------------------ */

struct FftPlan {
    int M = 0;
    std::vector<cplx> tw;      /* e^{+2 pi i m / M}, m < M/2 */
    std::vector<cplx> twist;   /* w^j, j < M */
    std::vector<int> rev;
};

static std::mutex g_plan_mu;
static std::map<int, FftPlan*> g_plans;

static cplx cis_pi(long double num, long double den) {
    /* e^{i pi num/den} from long-double trig, rounded once to double */
    long double ang = 3.14159265358979323846264338327950288L * num / den;
    return cplx((double)cosl(ang), (double)sinl(ang));
}

static const FftPlan* plan_for(int N) {
    std::lock_guard<std::mutex> g(g_plan_mu);
    auto it = g_plans.find(N);
    if (it != g_plans.end()) return it->second;
    FftPlan* pl = new FftPlan();
    const int M = N / 2;
    pl->M = M;
    pl->tw.resize(M / 2);
    for (int m = 0; m < M / 2; ++m) pl->tw[m] = cis_pi(2.0L * m, (long double)M);
    pl->twist.resize(M);
    for (int j = 0; j < M; ++j) pl->twist[j] = cis_pi((long double)j, (long double)N);
    pl->rev.resize(M);
    int bits = 0;
    while ((1 << bits) < M) ++bits;
    for (int i = 0; i < M; ++i) {
        int r = 0;
        for (int b = 0; b < bits; ++b) if (i & (1 << b)) r |= 1 << (bits - 1 - b);
        pl->rev[i] = r;
    }
    g_plans[N] = pl;
    return pl;
}

/* iterative radix-2 DIT; sign = +1 computes sum x_j e^{+2 pi i jk/M}, -1 the conjugate kernel */
static void fft_core(const FftPlan* pl, cplx* x, int sign) {
    const int M = pl->M;
    for (int i = 0; i < M; ++i) if (i < pl->rev[i]) std::swap(x[i], x[pl->rev[i]]);
    const bool fault = (g_fault == 5);
    for (int len = 2; len <= M; len <<= 1) {
        const int half = len >> 1, step = M / len;
        for (int s = 0; s < M; s += len)
            for (int j = 0; j < half; ++j) {
                cplx w = pl->tw[j * step];
                if (sign < 0 || fault) w = std::conj(w);
                cplx u = x[s + j], v = x[s + j + half] * w;
                x[s + j] = u + v;
                x[s + j + half] = u - v;
            }
    }
}

static void fft_forward_int(const FftPlan* pl, cplx* F, const int32_t* p) {
    const int M = pl->M;
    for (int j = 0; j < M; ++j) F[j] = cplx((double)p[j], (double)p[j + M]) * pl->twist[j];
    fft_core(pl, F, +1);
}

/* inverse to integer-unit doubles: out[j] = p_j (real, unreduced) for j < N */
static void fft_inverse_real(const FftPlan* pl, double* out, cplx* F) {
    const int M = pl->M;
    fft_core(pl, F, -1);
    const double s = 1.0 / M;
    for (int j = 0; j < M; ++j) {
        cplx z = F[j] * std::conj(pl->twist[j]) * s;
        out[j] = z.real();
        out[j + M] = z.imag();
    }
}

extern "C" void orc_fft_forward(double* out, const int32_t* poly, int32_t N) {
    const FftPlan* pl = plan_for(N);
    std::vector<cplx> F(pl->M);
    fft_forward_int(pl, F.data(), poly);
    for (int i = 0; i < pl->M; ++i) { out[2 * i] = F[i].real(); out[2 * i + 1] = F[i].imag(); }
}

extern "C" void orc_fft_inverse_raw(double* out, const double* in, int32_t N) {
    const FftPlan* pl = plan_for(N);
    std::vector<cplx> F(pl->M);
    for (int i = 0; i < pl->M; ++i) F[i] = cplx(in[2 * i], in[2 * i + 1]);
    fft_inverse_real(pl, out, F.data());
}

/* The reference's convention, tGswFFTExternMulToTLwe.cpp:97-130: a 2N-point DFT with
 * kernel e^{-2 pi i jk/2N} (twidI.txt sign) of [p/2, -p/2], conjugated, odd outputs:
 * LagrangeHalfC[i] = p(w^(2i+1)), i < N/2.  Restated by direct summation (O(N^2)), independent
 * of the folded transform above, so the tests can pin the permutation between the two. */
extern "C" void orc_lagrange_half_reference(double* out, const int32_t* poly, int32_t N) {
    for (int i = 0; i < N / 2; ++i) {
        long double re = 0, im = 0;
        for (int j = 0; j < N; ++j) {
            /* X^j at w^(2i+1): angle pi*j*(2i+1)/N, reduced mod 2N to keep the argument small */
            long long e = ((long long)j * (2 * i + 1)) % (2LL * N);
            long double ang = 3.14159265358979323846264338327950288L * e / N;
            re += poly[j] * cosl(ang);
            im += poly[j] * sinl(ang);
        }
        out[2 * i] = (double)re;
        out[2 * i + 1] = (double)im;
    }
}

/* TorusPolynomial_fft.cpp:4-43, the reference's way back from the FFT domain: the N/2 values p_i
 * go to the odd slots 2i+1 of a 2N-point vector and their conjugates to slots 2N-1-2i, a 2N-point
 * DFT with kernel e^{-2 pi i st/2N} (twidI.txt sign) is taken, and coefficient t is
 * int32(int64(Re(out_t) / N * 2^32)) -- truncation toward zero, as there.  Restated by direct
 * summation; `in` holds N/2 complex values (re, im interleaved). */
extern "C" void orc_torus_from_lagrange_reference(int32_t* out, const double* in, int32_t N) {
    const long double pi = 3.14159265358979323846264338327950288L;
    for (int t = 0; t < N; ++t) {
        long double re = 0;
        for (int i = 0; i < N / 2; ++i) {
            const long double pr = in[2 * i], pi_ = in[2 * i + 1];
            /* slot s = 2i+1 holds p_i, slot 2N-1-2i holds conj(p_i); angle -2 pi s t / 2N, mod 2N */
            const long long s1 = ((long long)(2 * i + 1) * t) % (2LL * N);
            const long long s2 = ((long long)(2 * N - 1 - 2 * i) * t) % (2LL * N);
            const long double a1 = -pi * s1 / N, a2 = -pi * s2 / N;
            re += pr * cosl(a1) - pi_ * sinl(a1);     /* Re(p e^{i a1}) */
            re += pr * cosl(a2) + pi_ * sinl(a2);     /* Re(conj(p) e^{i a2}) */
        }
        const long double x = re / N * 4294967296.0L;
        out[t] = (int32_t)(uint32_t)(int64_t)x;
    }
}

/* ------------------------------------------------------------------------ */
/* external product                                                         */
/* ------------------------------------------------------------------------ */

/* tGswFFTExternMulToTLwe.cpp:203-247 with the FFT replaced by exact int64 arithmetic:
 * acc <- sum over rows r = q*l + lvl of decomp_lvl(acc_q) * tgsw[r].  raw (optional) receives
 * the unreduced int64 sums [(k+1)][N]; acc receives them mod 2^32. */
extern "C" void orc_extern_product_exact(const orc_params* p, int32_t* acc, const int32_t* tgsw, int64_t* raw) {
    const int N = p->N, k = p->k, l = p->l;
    std::vector<int32_t> dec((size_t)l * N);
    std::vector<int64_t> sum((size_t)(k + 1) * N, 0), prod(N);
    for (int q = 0; q <= k; ++q) {
        orc_decompose(dec.data(), acc + (size_t)q * N, N, l, p->Bgbit);
        for (int lvl = 0; lvl < l; ++lvl) {
            const int r = q * l + lvl;
            for (int qq = 0; qq <= k; ++qq) {
                const int src_q = (g_fault == 6) ? (k - qq) : qq;
                const int32_t* kp = tgsw + ((size_t)r * (k + 1) + src_q) * N;
                orc_negacyclic_mul_i64(prod.data(), dec.data() + (size_t)lvl * N, kp, N);
                for (int j = 0; j < N; ++j) sum[(size_t)qq * N + j] += prod[j];
            }
        }
    }
    for (size_t i = 0; i < sum.size(); ++i) {
        if (raw) raw[i] = sum[i];
        acc[i] = (int32_t)(uint32_t)(uint64_t)sum[i];
    }
}

/* FFT-domain bootstrapping key: for each (i, row, q) the folded transform of the torus polynomial */
struct FftKey {
    orc_params p;
    std::vector<cplx> data;  /* [n][(k+1)l][(k+1)][N/2] */
};

extern "C" void* orc_fft_key_create(const orc_params* p, const int32_t* bk) {
    FftKey* fk = new FftKey();
    fk->p = *p;
    const int N = p->N, M = N / 2;
    const size_t polys = (size_t)p->n * (p->k + 1) * p->l * (p->k + 1);
    fk->data.resize(polys * M);
    const FftPlan* pl = plan_for(N);
#pragma omp parallel for schedule(static)
    for (long long q = 0; q < (long long)polys; ++q) fft_forward_int(pl, fk->data.data() + q * M, bk + q * N);
    return fk;
}

extern "C" void orc_fft_key_free(void* h) { delete (FftKey*)h; }

/* One external product through the FP64 FFT path.  Conversion back to torus follows
 * TorusPolynomial_fft.cpp:39-41 when truncate != 0 (int64 truncation of the scaled real part);
 * otherwise rounds to nearest, which reproduces the exact int64 product whenever the FFT error
 * is below 1/2. */
static void extern_product_fft(const orc_params* p, int32_t* acc, const cplx* keyrows, double* raw, int truncate,
                               std::vector<int32_t>& dec, std::vector<cplx>& D, std::vector<cplx>& S,
                               std::vector<double>& out) {
    const int N = p->N, k = p->k, l = p->l, M = N / 2;
    const FftPlan* pl = plan_for(N);
    for (auto& s : S) s = 0;
    for (int q = 0; q <= k; ++q) {
        orc_decompose(dec.data(), acc + (size_t)q * N, N, l, p->Bgbit);
        for (int lvl = 0; lvl < l; ++lvl) {
            fft_forward_int(pl, D.data(), dec.data() + (size_t)lvl * N);
            const int r = q * l + lvl;
            for (int qq = 0; qq <= k; ++qq) {
                const int src_q = (g_fault == 6) ? (k - qq) : qq;
                const cplx* kp = keyrows + ((size_t)r * (k + 1) + src_q) * M;
                cplx* sp = S.data() + (size_t)qq * M;
                for (int m = 0; m < M; ++m) sp[m] += D[m] * kp[m];
            }
        }
    }
    for (int qq = 0; qq <= k; ++qq) {
        fft_inverse_real(pl, out.data(), S.data() + (size_t)qq * M);
        for (int j = 0; j < N; ++j) {
            const double v = out[j];
            if (raw) raw[(size_t)qq * N + j] = v;
            int64_t iv = truncate ? (int64_t)v : (int64_t)std::llrint(v);
            acc[(size_t)qq * N + j] = (int32_t)(uint32_t)(uint64_t)iv;
        }
    }
}

struct Scratch {
    std::vector<int32_t> dec, tmp, tmp2;
    std::vector<cplx> D, S;
    std::vector<double> out;
    Scratch(const orc_params* p)
        : dec((size_t)p->l * p->N), tmp((size_t)(p->k + 1) * p->N), tmp2((size_t)(p->k + 1) * p->N),
          D(p->N / 2), S((size_t)(p->k + 1) * (p->N / 2)), out(p->N) {}
};

extern "C" void orc_extern_product_fft(const orc_params* p, int32_t* acc, const void* fkey, int32_t row_set,
                                       double* raw, int32_t truncate) {
    const FftKey* fk = (const FftKey*)fkey;
    const size_t per = (size_t)(p->k + 1) * p->l * (p->k + 1) * (p->N / 2);
    Scratch s(p);
    extern_product_fft(p, acc, fk->data.data() + (size_t)row_set * per, raw, truncate, s.dec, s.D, s.S, s.out);
}

/* ------------------------------------------------------------------------ */
/* blind rotation, extraction, keyswitch, bootstrapping                     */
/* ------------------------------------------------------------------------ */

/* tfhe_blindRotate_FFT.cpp:112-127 with the loop bound n (the device code stops at 1, :112),
 * loop body = tfhe_MuxRotate_FFT.cpp:56-85: t = (X^a - 1) acc; t = bk_i [.] t; acc = t + acc.
 * Uses the exact int64 external product when fkey == NULL, else the FP64 FFT path. */
static void blind_rotate(const orc_params* p, int32_t* acc, const int32_t* bara, const int32_t* bk,
                         const FftKey* fk, Scratch& s) {
    const int N = p->N, k = p->k;
    const size_t tgsw_sz = (size_t)(k + 1) * p->l * (k + 1) * N;
    const size_t per = (size_t)(k + 1) * p->l * (k + 1) * (N / 2);
    for (int i = 0; i < p->n; ++i) {
        const int32_t a = bara[i];
        if (a == 0) continue;                                         /* :114 */
        for (int q = 0; q <= k; ++q)
            orc_poly_mul_xai_minus_one(s.tmp.data() + (size_t)q * N, a, acc + (size_t)q * N, N);
        if (fk) extern_product_fft(p, s.tmp.data(), fk->data.data() + i * per, nullptr, 0, s.dec, s.D, s.S, s.out);
        else orc_extern_product_exact(p, s.tmp.data(), bk + i * tgsw_sz, nullptr);
        for (size_t j = 0; j < (size_t)(k + 1) * N; ++j)
            acc[j] = (int32_t)((uint32_t)acc[j] + (uint32_t)s.tmp[j]);  /* tLweAddTo, tfhe_MuxRotate_FFT.cpp:48-54 */
    }
}

extern "C" void orc_blind_rotate(const orc_params* p, int32_t* acc, const int32_t* bara, const int32_t* bk,
                                 const void* fkey) {
    Scratch s(p);
    blind_rotate(p, acc, bara, bk, (const FftKey*)fkey, s);
}

/* tfhe_bootstrap_woKS_FFT.cpp:63-66 (testvect = [mu]*N), mulByXaiOrCopy.cpp:25-41 (X^{2N-barb}),
 * tb :317 (tLweNoiselessTrivial: a = 0, b = v). */
extern "C" void orc_test_vector(const orc_params* p, int32_t* acc, int32_t barb, int32_t mu) {
    const int N = p->N, k = p->k;
    std::vector<int32_t> tv(N, mu);
    for (size_t j = 0; j < (size_t)k * N; ++j) acc[j] = 0;
    int32_t* b = acc + (size_t)k * N;
    if (barb != 0) orc_poly_mul_xai(b, 2 * N - barb, tv.data(), N);
    else for (int j = 0; j < N; ++j) b[j] = tv[j];
}

/* upstream tLweExtractLweSample at index 0 (absent from the reference, which never extracts;
 * pinned by the closed form below and by the LWE phase = TLWE phase coefficient 0 identity):
 * u.a[q*N + 0] = acc.a[q][0], u.a[q*N + j] = -acc.a[q][N-j] (j >= 1), u.b = acc.b[0]. */
extern "C" void orc_extract(const orc_params* p, int32_t* u, const int32_t* acc) {
    const int N = p->N, k = p->k;
    const int shift = (g_fault == 3) ? 1 : 0;
    for (int q = 0; q < k; ++q) {
        const int32_t* a = acc + (size_t)q * N;
        u[(size_t)q * N] = a[0];
        for (int j = 1; j < N; ++j) {
            int src = (N - j + shift) % N;
            u[(size_t)q * N + j] = (int32_t)(0u - (uint32_t)a[src]);
        }
    }
    u[(size_t)k * N] = acc[(size_t)k * N];
}

/* lwe-keyswitch-functions.cpp:228-241 and :114-129 */
extern "C" void orc_keyswitch(const orc_params* p, int32_t* res, const int32_t* ksk, const int32_t* u) {
    const int n = p->n, kN = p->k * p->N, t = p->ks_t, basebit = p->ks_basebit;
    const int base = 1 << basebit;
    const uint32_t prec_offset = 1u << (32 - (1 + basebit * t));
    const uint32_t mask = (uint32_t)base - 1;
    std::vector<uint32_t> r(n + 1, 0);
    r[n] = (uint32_t)u[kN];                                          /* lweNoiselessTrivial(res, b) */
    const int tt = (g_fault == 4) ? t - 1 : t;
    for (int i = 0; i < kN; ++i) {
        const uint32_t aibar = (uint32_t)u[i] + prec_offset;
        for (int j = 0; j < tt; ++j) {
            const uint32_t aij = (aibar >> (32 - (j + 1) * basebit)) & mask;
            if (aij != 0) {
                const int32_t* row = ksk + (((size_t)i * t + j) * base + aij) * (size_t)(n + 1);
                for (int c = 0; c <= n; ++c) r[c] -= (uint32_t)row[c];
            }
        }
    }
    for (int c = 0; c <= n; ++c) res[c] = (int32_t)r[c];
}

static void bootstrap_woks(const orc_params* p, int32_t* u, const int32_t* x, int32_t mu, const int32_t* bk,
                           const FftKey* fk, Scratch& s) {
    const int n = p->n, N = p->N, k = p->k;
    const int Nx2 = 2 * N;
    std::vector<int32_t> bara(n);
    int32_t barb = orc_modswitch_from_torus32(x[n], Nx2);            /* woKS :57 */
    for (int i = 0; i < n; ++i) bara[i] = orc_modswitch_from_torus32(x[i], Nx2);  /* :58-61 */
    std::vector<int32_t> acc((size_t)(k + 1) * N);
    orc_test_vector(p, acc.data(), barb, mu);
    blind_rotate(p, acc.data(), bara.data(), bk, fk, s);
    orc_extract(p, u, acc.data());
}

extern "C" void orc_bootstrap_woks(const orc_params* p, int32_t* u, const int32_t* x, int32_t mu,
                                   const int32_t* bk, const void* fkey) {
    Scratch s(p);
    bootstrap_woks(p, u, x, mu, bk, (const FftKey*)fkey, s);
}

extern "C" void orc_bootstrap(const orc_params* p, int32_t* out, const int32_t* x, int32_t mu,
                              const int32_t* bk, const void* fkey, const int32_t* ksk) {
    Scratch s(p);
    std::vector<int32_t> u((size_t)p->k * p->N + 1);
    bootstrap_woks(p, u.data(), x, mu, bk, (const FftKey*)fkey, s);
    orc_keyswitch(p, out, ksk, u.data());
}

/* ------------------------------------------------------------------------ */
/* gates                                                                    */
/* ------------------------------------------------------------------------ */

/* Linear pre-combination tmp = (0, C) + alpha*a + beta*b.  OR is pinned by
 * tfhe_blindRotate_FFT_tb.cpp:273-276; the others are the upstream gate constants. */
static void gate_coeffs(int gate, int32_t* C, int32_t* alpha, int32_t* beta) {
    const int32_t e8 = orc_modswitch_to_torus32(1, 8), e4 = orc_modswitch_to_torus32(1, 4);
    switch (gate) {
        case ORC_GATE_NAND:  *C = e8;  *alpha = -1; *beta = -1; break;
        case ORC_GATE_AND:   *C = -e8; *alpha = 1;  *beta = 1;  break;
        case ORC_GATE_OR:    *C = e8;  *alpha = 1;  *beta = 1;  break;
        case ORC_GATE_XOR:   *C = e4;  *alpha = 2;  *beta = 2;  break;
        case ORC_GATE_XNOR:  *C = -e4; *alpha = -2; *beta = -2; break;
        case ORC_GATE_NOR:   *C = -e8; *alpha = -1; *beta = -1; break;
        case ORC_GATE_ANDNY: *C = -e8; *alpha = -1; *beta = 1;  break;
        case ORC_GATE_ANDYN: *C = -e8; *alpha = 1;  *beta = -1; break;
        case ORC_GATE_ORNY:  *C = e8;  *alpha = -1; *beta = 1;  break;
        case ORC_GATE_ORYN:  *C = e8;  *alpha = 1;  *beta = -1; break;
        default:             *C = 0;   *alpha = 0;  *beta = 0;  break;
    }
}

extern "C" void orc_gate_combine(const orc_params* p, int32_t gate, int32_t* tmp, const int32_t* a, const int32_t* b) {
    int32_t C, al, be;
    gate_coeffs(gate, &C, &al, &be);
    const int n = p->n;
    for (int i = 0; i < n; ++i)
        tmp[i] = (int32_t)((uint32_t)al * (uint32_t)a[i] + (uint32_t)be * (uint32_t)b[i]);
    tmp[n] = (int32_t)((uint32_t)C + (uint32_t)al * (uint32_t)a[n] + (uint32_t)be * (uint32_t)b[n]);
}

static void gate_one(const orc_params* p, int gate, int32_t* out, const int32_t* a, const int32_t* b,
                     const int32_t* c, const int32_t* bk, const FftKey* fk, const int32_t* ksk, Scratch& s) {
    const int n = p->n, kN = p->k * p->N;
    const int32_t MU = orc_modswitch_to_torus32(1, 8);
    std::vector<int32_t> tmp(n + 1);
    if (gate != ORC_GATE_MUX) {
        orc_gate_combine(p, gate, tmp.data(), a, b);
        std::vector<int32_t> u((size_t)kN + 1);
        bootstrap_woks(p, u.data(), tmp.data(), MU, bk, fk, s);
        orc_keyswitch(p, out, ksk, u.data());
        return;
    }
    /* upstream bootsMUX: u1 = woKS(AND(a,b)), u2 = woKS(AND(not a, c)), KS((0,1/8) + u1 + u2) */
    std::vector<int32_t> u1((size_t)kN + 1), u2((size_t)kN + 1);
    orc_gate_combine(p, ORC_GATE_AND, tmp.data(), a, b);
    bootstrap_woks(p, u1.data(), tmp.data(), MU, bk, fk, s);
    orc_gate_combine(p, ORC_GATE_ANDNY, tmp.data(), a, c);
    bootstrap_woks(p, u2.data(), tmp.data(), MU, bk, fk, s);
    for (int i = 0; i <= kN; ++i) u1[i] = (int32_t)((uint32_t)u1[i] + (uint32_t)u2[i]);
    u1[kN] = (int32_t)((uint32_t)u1[kN] + (uint32_t)MU);
    orc_keyswitch(p, out, ksk, u1.data());
}

extern "C" void orc_gate(const orc_params* p, int32_t gate, int32_t* out, const int32_t* a, const int32_t* b,
                         const int32_t* c, const int32_t* bk, const void* fkey, const int32_t* ksk) {
    Scratch s(p);
    gate_one(p, gate, out, a, b, c, bk, (const FftKey*)fkey, ksk, s);
}

/* CPU baseline: `count` independent gates over the FP64 FFT path, one gate per OpenMP thread.
 * Returns wall seconds. */
extern "C" double orc_gate_batch(const orc_params* p, int32_t gate, int32_t* out, const int32_t* a,
                                 const int32_t* b, const int32_t* c, int32_t count, const void* fkey,
                                 const int32_t* ksk, int32_t nthreads) {
    const size_t st = (size_t)p->n + 1;
    auto t0 = std::chrono::steady_clock::now();
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
    {
        Scratch s(p);
#pragma omp for schedule(dynamic, 1)
        for (int g = 0; g < count; ++g)
            gate_one(p, gate, out + g * st, a + g * st, b ? b + g * st : nullptr, c ? c + g * st : nullptr,
                     nullptr, (const FftKey*)fkey, ksk, s);
    }
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}
