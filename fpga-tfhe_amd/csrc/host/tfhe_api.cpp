// tfhe_api.cpp -- upstream-compatible host API of libtfhe_mi355x: parameters, key generation,
// encryption, LWE arithmetic and the boots* gates (which run on the GPU through the flat ABI).
//
// Draw order and arithmetic follow the reference's copies of the upstream library:
//   kernel/src/numeric-functions.cpp      global engine, seed, gaussian32, dtot32, modSwitch
//   kernel/src/lwe-functions.cpp          lweKeyGen, lweSymEncrypt, lwePhase, LWE arithmetic
//   kernel/src/lwe-keyswitch-functions.cpp lweCreateKeySwitchKey (recentred noise), lweKeySwitch
// TLWE/TGSW key generation and the gate constants other than OR are upstream-only (the reference
// calls them from test_bench/ and tfhe_blindRotate_FFT_tb.cpp:159-325); they follow the CGGI
// construction those call sites rely on.
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "internal.h"

using namespace tfhe;

// ------------------------------------------------------------------------------------------
// torus numerics (numeric-functions.cpp:11-76)
// ------------------------------------------------------------------------------------------
static std::default_random_engine g_generator;
static std::uniform_int_distribution<int32_t> g_uniform_torus(INT32_MIN, INT32_MAX);

extern "C" void tfhe_random_generator_setSeed(uint32_t* values, int32_t size) {
    std::seed_seq seeds(values, values + size);
    g_generator.seed(seeds);
}

extern "C" Torus32 dtot32(double d) {
    const double two32 = 4294967296.0;
    return (int32_t)(int64_t)((d - (double)(int64_t)d) * two32);
}

extern "C" double t32tod(Torus32 x) { return (double)x / 4294967296.0; }

extern "C" Torus32 gaussian32(Torus32 message, double sigma) {
    std::normal_distribution<double> distribution(0., sigma);
    const double err = distribution(g_generator);
    return (int32_t)((uint32_t)message + (uint32_t)dtot32(err));
}

extern "C" Torus32 approxPhase(Torus32 phase, int32_t Msize) {
    const uint64_t interv = ((UINT64_C(1) << 63) / (uint64_t)Msize) * 2;
    const uint64_t half_interval = interv / 2;
    uint64_t phase64 = ((uint64_t)(uint32_t)phase << 32) + half_interval;
    phase64 -= phase64 % interv;
    return (int32_t)(uint32_t)(phase64 >> 32);
}

extern "C" int32_t modSwitchFromTorus32(Torus32 phase, int32_t Msize) {
    const uint64_t interv = ((UINT64_C(1) << 63) / (uint64_t)Msize) * 2;
    const uint64_t half_interval = interv / 2;
    const uint64_t phase64 = ((uint64_t)(uint32_t)phase << 32) + half_interval;
    return (int32_t)(phase64 / interv);
}

extern "C" Torus32 modSwitchToTorus32(int32_t mu, int32_t Msize) {
    const uint64_t interv = ((UINT64_C(1) << 63) / (uint64_t)Msize) * 2;
    const uint64_t phase64 = (uint64_t)(int64_t)mu * interv;
    return (int32_t)(uint32_t)(phase64 >> 32);
}

static int32_t uniform_bit() {
    std::uniform_int_distribution<int32_t> distribution(0, 1);
    return distribution(g_generator);
}

// ------------------------------------------------------------------------------------------
// parameters
// ------------------------------------------------------------------------------------------
extern "C" LweParams* new_LweParams(int32_t n, double alpha_min, double alpha_max) {
    return new LweParams{n, alpha_min, alpha_max};
}
extern "C" void delete_LweParams(LweParams* obj) { delete obj; }

extern "C" TLweParams* new_TLweParams(int32_t N, int32_t k, double alpha_min, double alpha_max) {
    return new TLweParams{N, k, alpha_min, alpha_max, LweParams{k * N, alpha_min, alpha_max}};
}
extern "C" void delete_TLweParams(TLweParams* obj) { delete obj; }

extern "C" TGswParams* new_TGswParams(int32_t l, int32_t Bgbit, const TLweParams* tlwe_params) {
    TGswParams* p = new TGswParams();
    p->l = l;
    p->Bgbit = Bgbit;
    p->Bg = 1 << Bgbit;
    p->halfBg = p->Bg / 2;
    p->maskMod = (uint32_t)p->Bg - 1;
    p->tlwe_params = tlwe_params;
    p->kpl = (tlwe_params->k + 1) * l;
    p->h = new Torus32[l];
    uint32_t off = 0;
    for (int i = 0; i < l; ++i) {
        const uint32_t hi = 1u << (32 - (i + 1) * Bgbit);
        p->h[i] = (int32_t)hi;
        off += hi;
    }
    p->offset = off * (uint32_t)p->halfBg;  // 2164391936 for l=3, Bgbit=7 (constants.h:24)
    return p;
}
extern "C" void delete_TGswParams(TGswParams* obj) {
    if (!obj) return;
    delete[] obj->h;
    delete obj;
}

extern "C" TFheGateBootstrappingParameterSet* new_TFheGateBootstrappingParameterSet(
    int32_t ks_t, int32_t ks_basebit, const LweParams* in_out_params, const TGswParams* tgsw_params) {
    return new TFheGateBootstrappingParameterSet{ks_t, ks_basebit, in_out_params, tgsw_params};
}
extern "C" void delete_TFheGateBootstrappingParameterSet(TFheGateBootstrappingParameterSet* params) { delete params; }

static TFheGateBootstrappingParameterSet* make_params(int N, int k, int n, int l, int Bgbit, int ks_basebit,
                                                      int ks_t, double ks_stdev, double bk_stdev, double max_stdev) {
    LweParams* in_out = new_LweParams(n, ks_stdev, max_stdev);
    TLweParams* accum = new_TLweParams(N, k, bk_stdev, max_stdev);
    TGswParams* bkp = new_TGswParams(l, Bgbit, accum);
    return new_TFheGateBootstrappingParameterSet(ks_t, ks_basebit, in_out, bkp);
}

// kernel/src/include/constants.h:26-36 (the 128-bit branch, the one the reference builds)
extern "C" TFheGateBootstrappingParameterSet* new_default_gate_bootstrapping_parameters(int32_t minimum_lambda) {
    if (minimum_lambda > 128) fprintf(stderr, "tfhe_mi355x: parameters target about 128 bits of security\n");
    return make_params(1024, 1, 630, 3, 7, 2, 8, std::pow(2., -15), std::pow(2., -25), 0.012467);
}

extern "C" TFheGateBootstrappingParameterSet* new_gate_bootstrapping_parameters_N2048(void) {
    return make_params(2048, 1, 750, 3, 8, 2, 10, std::pow(2., -17), std::pow(2., -30), 0.012467);
}

extern "C" void delete_gate_bootstrapping_parameters(TFheGateBootstrappingParameterSet* params) {
    if (!params) return;
    const TGswParams* bkp = params->tgsw_params;
    delete_TLweParams(const_cast<TLweParams*>(bkp->tlwe_params));
    delete_TGswParams(const_cast<TGswParams*>(bkp));
    delete_LweParams(const_cast<LweParams*>(params->in_out_params));
    delete params;
}

// ------------------------------------------------------------------------------------------
// samples and polynomials
// ------------------------------------------------------------------------------------------
static void init_lwe(LweSample* s, int n) {
    s->a = new Torus32[n]();
    s->b = 0;
    s->current_variance = 0.;
}

extern "C" LweSample* new_LweSample(const LweParams* params) {
    LweSample* s = new LweSample;
    init_lwe(s, params->n);
    return s;
}
extern "C" LweSample* new_LweSample_array(int32_t nbelts, const LweParams* params) {
    LweSample* s = new LweSample[nbelts];
    for (int i = 0; i < nbelts; ++i) init_lwe(s + i, params->n);
    return s;
}
extern "C" void delete_LweSample(LweSample* obj) {
    if (!obj) return;
    delete[] obj->a;
    delete obj;
}
extern "C" void delete_LweSample_array(int32_t nbelts, LweSample* obj) {
    if (!obj) return;
    for (int i = 0; i < nbelts; ++i) delete[] obj[i].a;
    delete[] obj;
}
extern "C" LweSample* new_gate_bootstrapping_ciphertext(const TFheGateBootstrappingParameterSet* params) {
    return new_LweSample(params->in_out_params);
}
extern "C" LweSample* new_gate_bootstrapping_ciphertext_array(int32_t nbelems,
                                                              const TFheGateBootstrappingParameterSet* params) {
    return new_LweSample_array(nbelems, params->in_out_params);
}
extern "C" void delete_gate_bootstrapping_ciphertext(LweSample* sample) { delete_LweSample(sample); }
extern "C" void delete_gate_bootstrapping_ciphertext_array(int32_t nbelems, LweSample* samples) {
    delete_LweSample_array(nbelems, samples);
}

extern "C" TorusPolynomial* new_TorusPolynomial(int32_t N) {
    return new TorusPolynomial{N, new Torus32[N]()};
}
extern "C" void delete_TorusPolynomial(TorusPolynomial* obj) {
    if (!obj) return;
    delete[] obj->coefsT;
    delete obj;
}

static void init_tlwe(TLweSample* s, const TLweParams* p) {
    s->k = p->k;
    s->a = new TorusPolynomial[p->k + 1];
    for (int i = 0; i <= p->k; ++i) s->a[i] = TorusPolynomial{p->N, new Torus32[p->N]()};
    s->b = s->a + p->k;
    s->current_variance = 0.;
}
static void destroy_tlwe(TLweSample* s) {
    for (int i = 0; i <= s->k; ++i) delete[] s->a[i].coefsT;
    delete[] s->a;
}
extern "C" TLweSample* new_TLweSample(const TLweParams* params) {
    TLweSample* s = new TLweSample;
    init_tlwe(s, params);
    return s;
}
extern "C" void delete_TLweSample(TLweSample* obj) {
    if (!obj) return;
    destroy_tlwe(obj);
    delete obj;
}

// ------------------------------------------------------------------------------------------
// LWE (lwe-functions.cpp)
// ------------------------------------------------------------------------------------------
extern "C" void lweKeyGen(LweKey* result) {
    const int n = result->params->n;
    for (int i = 0; i < n; ++i) result->key[i] = uniform_bit();
}

extern "C" void lweSymEncrypt(LweSample* result, Torus32 message, double alpha, const LweKey* key) {
    const int n = key->params->n;
    uint32_t b = (uint32_t)gaussian32(message, alpha);
    for (int i = 0; i < n; ++i) {
        result->a[i] = g_uniform_torus(g_generator);
        b += (uint32_t)result->a[i] * (uint32_t)key->key[i];
    }
    result->b = (int32_t)b;
    result->current_variance = alpha * alpha;
}

static void lwe_encrypt_external_noise(LweSample* result, Torus32 message, double noise, double alpha,
                                       const LweKey* key) {
    const int n = key->params->n;
    uint32_t b = (uint32_t)message + (uint32_t)dtot32(noise);
    for (int i = 0; i < n; ++i) {
        result->a[i] = g_uniform_torus(g_generator);
        b += (uint32_t)result->a[i] * (uint32_t)key->key[i];
    }
    result->b = (int32_t)b;
    result->current_variance = alpha * alpha;
}

extern "C" Torus32 lwePhase(const LweSample* sample, const LweKey* key) {
    const int n = key->params->n;
    uint32_t axs = 0;
    for (int i = 0; i < n; ++i) axs += (uint32_t)sample->a[i] * (uint32_t)key->key[i];
    return (int32_t)((uint32_t)sample->b - axs);
}

extern "C" Torus32 lweSymDecrypt(const LweSample* sample, const LweKey* key, int32_t Msize) {
    return approxPhase(lwePhase(sample, key), Msize);
}

extern "C" void lweClear(LweSample* result, const LweParams* params) {
    for (int i = 0; i < params->n; ++i) result->a[i] = 0;
    result->b = 0;
    result->current_variance = 0.;
}
extern "C" void lweCopy(LweSample* result, const LweSample* sample, const LweParams* params) {
    if (result == sample) return;
    for (int i = 0; i < params->n; ++i) result->a[i] = sample->a[i];
    result->b = sample->b;
    result->current_variance = sample->current_variance;
}
extern "C" void lweNegate(LweSample* result, const LweSample* sample, const LweParams* params) {
    for (int i = 0; i < params->n; ++i) result->a[i] = (int32_t)(0u - (uint32_t)sample->a[i]);
    result->b = (int32_t)(0u - (uint32_t)sample->b);
    result->current_variance = sample->current_variance;
}
extern "C" void lweNoiselessTrivial(LweSample* result, Torus32 mu, const LweParams* params) {
    for (int i = 0; i < params->n; ++i) result->a[i] = 0;
    result->b = mu;
    result->current_variance = 0.;
}
extern "C" void lweAddTo(LweSample* result, const LweSample* sample, const LweParams* params) {
    for (int i = 0; i < params->n; ++i) result->a[i] = (int32_t)((uint32_t)result->a[i] + (uint32_t)sample->a[i]);
    result->b = (int32_t)((uint32_t)result->b + (uint32_t)sample->b);
    result->current_variance += sample->current_variance;
}
extern "C" void lweSubTo(LweSample* result, const LweSample* sample, const LweParams* params) {
    for (int i = 0; i < params->n; ++i) result->a[i] = (int32_t)((uint32_t)result->a[i] - (uint32_t)sample->a[i]);
    result->b = (int32_t)((uint32_t)result->b - (uint32_t)sample->b);
    result->current_variance += sample->current_variance;
}
extern "C" void lweAddMulTo(LweSample* result, int32_t p, const LweSample* sample, const LweParams* params) {
    for (int i = 0; i < params->n; ++i)
        result->a[i] = (int32_t)((uint32_t)result->a[i] + (uint32_t)p * (uint32_t)sample->a[i]);
    result->b = (int32_t)((uint32_t)result->b + (uint32_t)p * (uint32_t)sample->b);
    result->current_variance += (double)p * p * sample->current_variance;
}
extern "C" void lweSubMulTo(LweSample* result, int32_t p, const LweSample* sample, const LweParams* params) {
    for (int i = 0; i < params->n; ++i)
        result->a[i] = (int32_t)((uint32_t)result->a[i] - (uint32_t)p * (uint32_t)sample->a[i]);
    result->b = (int32_t)((uint32_t)result->b - (uint32_t)p * (uint32_t)sample->b);
    result->current_variance += (double)p * p * sample->current_variance;
}

// ------------------------------------------------------------------------------------------
// keys
// ------------------------------------------------------------------------------------------
static LweKey* new_LweKey(const LweParams* params) {
    return new LweKey{params, new int32_t[params->n]()};
}
static void delete_LweKey(LweKey* k) {
    if (!k) return;
    delete[] k->key;
    delete k;
}

static TGswKey* new_TGswKey(const TGswParams* params) {
    const TLweParams* tp = params->tlwe_params;
    TGswKey* k = new TGswKey();
    k->params = params;
    k->tlwe_params = tp;
    IntPolynomial* polys = new IntPolynomial[tp->k];
    for (int i = 0; i < tp->k; ++i) polys[i] = IntPolynomial{tp->N, new int32_t[tp->N]()};
    k->tlwe_key = TLweKey{tp, polys};
    k->key = polys;
    return k;
}
static void delete_TGswKey(TGswKey* k) {
    if (!k) return;
    for (int i = 0; i < k->tlwe_params->k; ++i) delete[] k->tlwe_key.key[i].coefs;
    delete[] k->tlwe_key.key;
    delete k;
}

static LweKeySwitchKey* new_LweKeySwitchKey(int32_t n, int32_t t, int32_t basebit, const LweParams* out_params) {
    LweKeySwitchKey* ks = new LweKeySwitchKey();
    ks->n = n;
    ks->t = t;
    ks->basebit = basebit;
    ks->base = 1 << basebit;
    ks->out_params = out_params;
    ks->ks0_raw = new_LweSample_array(n * t * ks->base, out_params);
    ks->ks1_raw = new LweSample*[(size_t)n * t];
    for (int i = 0; i < n * t; ++i) ks->ks1_raw[i] = ks->ks0_raw + (size_t)i * ks->base;
    ks->ks = new LweSample**[n];
    for (int i = 0; i < n; ++i) ks->ks[i] = ks->ks1_raw + (size_t)i * t;
    return ks;
}
static void delete_LweKeySwitchKey(LweKeySwitchKey* ks) {
    if (!ks) return;
    delete_LweSample_array(ks->n * ks->t * ks->base, ks->ks0_raw);
    delete[] ks->ks1_raw;
    delete[] ks->ks;
    delete ks;
}

static void init_tgsw(TGswSample* s, const TGswParams* p) {
    const int kpl = p->kpl, k = p->tlwe_params->k, l = p->l;
    s->k = k;
    s->l = l;
    s->all_sample = new TLweSample[kpl];
    for (int r = 0; r < kpl; ++r) init_tlwe(s->all_sample + r, p->tlwe_params);
    s->bloc_sample = new TLweSample*[k + 1];
    for (int b = 0; b <= k; ++b) s->bloc_sample[b] = s->all_sample + b * l;
}
static void destroy_tgsw(TGswSample* s, int kpl) {
    for (int r = 0; r < kpl; ++r) destroy_tlwe(s->all_sample + r);
    delete[] s->all_sample;
    delete[] s->bloc_sample;
}

static void init_tgsw_fft(TGswSampleFFT* s, const TGswParams* p, void* owner) {
    const int kpl = p->kpl, k = p->tlwe_params->k, l = p->l;
    s->k = k;
    s->l = l;
    s->all_samples = new TLweSampleFFT[kpl];
    for (int r = 0; r < kpl; ++r) {
        TLweSampleFFT* t = s->all_samples + r;
        t->k = k;
        t->a = new LagrangeHalfCPolynomial[k + 1];
        for (int q = 0; q <= k; ++q) t->a[q] = LagrangeHalfCPolynomial{nullptr, owner};
        t->b = t->a + k;
        t->current_variance = 0.;
    }
    s->sample = new TLweSampleFFT*[k + 1];
    for (int b = 0; b <= k; ++b) s->sample[b] = s->all_samples + b * l;
}
static void destroy_tgsw_fft(TGswSampleFFT* s, int kpl) {
    for (int r = 0; r < kpl; ++r) delete[] s->all_samples[r].a;
    delete[] s->all_samples;
    delete[] s->sample;
}

// negacyclic product with uint32 wrap: res += ip * tp (upstream torusPolynomialAddMulR)
static void poly_addmul(uint32_t* res, const int32_t* ip, const int32_t* tp, int N) {
    for (int i = 0; i < N; ++i) {
        const uint32_t a = (uint32_t)ip[i];
        if (a == 0) continue;
        for (int j = 0; j < N - i; ++j) res[i + j] += a * (uint32_t)tp[j];
        for (int j = N - i; j < N; ++j) res[i + j - N] -= a * (uint32_t)tp[j];
    }
}

// upstream tGswSymEncryptInt: every row = TLWE encryption of zero (tLweSymEncryptZero: the N noisy
// b coefficients are drawn first, then each mask polynomial), then message * h[i] added to
// coefficient 0 of polynomial `bloc` of row bloc*l + i (tGswAddMuIntH).
static void tgsw_encrypt_int(TGswSample* s, int32_t message, double alpha, const TGswKey* key) {
    const TGswParams* p = key->params;
    const int N = key->tlwe_params->N, k = key->tlwe_params->k;
    std::vector<uint32_t> acc(N);
    for (int r = 0; r < p->kpl; ++r) {
        TLweSample* row = s->all_sample + r;
        for (int j = 0; j < N; ++j) row->b->coefsT[j] = gaussian32(0, alpha);
        for (int q = 0; q < k; ++q) {
            for (int j = 0; j < N; ++j) row->a[q].coefsT[j] = g_uniform_torus(g_generator);
            for (int j = 0; j < N; ++j) acc[j] = (uint32_t)row->b->coefsT[j];
            poly_addmul(acc.data(), key->tlwe_key.key[q].coefs, row->a[q].coefsT, N);
            for (int j = 0; j < N; ++j) row->b->coefsT[j] = (int32_t)acc[j];
        }
        row->current_variance = alpha * alpha;
    }
    for (int b = 0; b <= k; ++b)
        for (int i = 0; i < p->l; ++i) {
            Torus32* c = s->bloc_sample[b][i].a[b].coefsT;
            c[0] = (int32_t)((uint32_t)c[0] + (uint32_t)message * (uint32_t)p->h[i]);
        }
}

// lwe-keyswitch-functions.cpp:163-215: recentred Gaussian noise vector drawn first, then each
// (i, j, h>=1) sample; h = 0 is a trivial zero that the keyswitch never reads.
static void create_ksk(LweKeySwitchKey* result, const int32_t* in_key, const LweKey* out_key) {
    const int n = result->n, t = result->t, basebit = result->basebit, base = 1 << basebit;
    const double alpha = out_key->params->alpha_min;
    const int sizeks = n * t * (base - 1);
    std::vector<double> noise(sizeks);
    double err = 0;
    for (int i = 0; i < sizeks; ++i) {
        std::normal_distribution<double> distribution(0., alpha);
        noise[i] = distribution(g_generator);
        err += noise[i];
    }
    err /= sizeks;
    for (int i = 0; i < sizeks; ++i) noise[i] -= err;
    int index = 0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < t; ++j) {
            lweNoiselessTrivial(&result->ks[i][j][0], 0, out_key->params);
            for (int h = 1; h < base; ++h) {
                const uint32_t mess = (uint32_t)in_key[i] * (uint32_t)h * (1u << (32 - (j + 1) * basebit));
                lwe_encrypt_external_noise(&result->ks[i][j][h], (int32_t)mess, noise[index], alpha, out_key);
                index += 1;
            }
        }
}

extern "C" TFheGateBootstrappingSecretKeySet* new_random_gate_bootstrapping_secret_keyset(
    const TFheGateBootstrappingParameterSet* params) {
    const LweParams* in_out = params->in_out_params;
    const TGswParams* bkp = params->tgsw_params;
    const TLweParams* accum = bkp->tlwe_params;
    const int n = in_out->n, N = accum->N, k = accum->k;

    LweKey* lwe_key = new_LweKey(in_out);
    lweKeyGen(lwe_key);
    TGswKey* tgsw_key = new_TGswKey(bkp);
    for (int q = 0; q < k; ++q)
        for (int j = 0; j < N; ++j) tgsw_key->tlwe_key.key[q].coefs[j] = uniform_bit();

    TFheGateBootstrappingCloudKeySet cloud = tfhe::alloc_cloud_keyset(params);
    LweBootstrappingKey* bk = const_cast<LweBootstrappingKey*>(cloud.bk);
    std::vector<int32_t> extracted((size_t)k * N);
    for (int q = 0; q < k; ++q)
        for (int j = 0; j < N; ++j) extracted[(size_t)q * N + j] = tgsw_key->tlwe_key.key[q].coefs[j];
    create_ksk(bk->ks, extracted.data(), lwe_key);
    for (int i = 0; i < n; ++i) tgsw_encrypt_int(bk->bk + i, lwe_key->key[i], accum->alpha_min, tgsw_key);

    TFheGateBootstrappingSecretKeySet* ks = new TFheGateBootstrappingSecretKeySet{params, lwe_key, tgsw_key, cloud};
    register_keyset(&ks->cloud);
    return ks;
}

namespace tfhe {

// The containers of a cloud key set (time-domain bk, keyswitch key, FFT-domain stubs), unfilled.
TFheGateBootstrappingCloudKeySet alloc_cloud_keyset(const TFheGateBootstrappingParameterSet* params) {
    const LweParams* in_out = params->in_out_params;
    const TGswParams* bkp = params->tgsw_params;
    const TLweParams* accum = bkp->tlwe_params;
    const int n = in_out->n, N = accum->N, k = accum->k;
    LweBootstrappingKey* bk = new LweBootstrappingKey();
    bk->in_out_params = in_out;
    bk->bk_params = bkp;
    bk->accum_params = accum;
    bk->extract_params = &accum->extracted_lweparams;
    bk->ks = new_LweKeySwitchKey(k * N, params->ks_t, params->ks_basebit, in_out);
    bk->bk = new TGswSample[n];
    for (int i = 0; i < n; ++i) init_tgsw(bk->bk + i, bkp);
    LweBootstrappingKeyFFT* bkfft = new LweBootstrappingKeyFFT();
    bkfft->in_out_params = in_out;
    bkfft->bk_params = bkp;
    bkfft->accum_params = accum;
    bkfft->extract_params = &accum->extracted_lweparams;
    bkfft->ks = bk->ks;
    TGswSampleFFT* samples = new TGswSampleFFT[n];
    for (int i = 0; i < n; ++i) init_tgsw_fft(samples + i, bkp, (void*)bkfft);
    bkfft->bkFFT = samples;
    return TFheGateBootstrappingCloudKeySet{params, bk, bkfft};
}

LweKey* alloc_lwe_key(const LweParams* params) { return new_LweKey(params); }
TGswKey* alloc_tgsw_key(const TGswParams* params) { return new_TGswKey(params); }
TFheGateBootstrappingParameterSet* alloc_params(const tfhe_flat_params& f, double lwe_alpha_min, double lwe_alpha_max,
                                                double tlwe_alpha_min, double tlwe_alpha_max) {
    LweParams* in_out = new_LweParams(f.n, lwe_alpha_min, lwe_alpha_max);
    TLweParams* accum = new_TLweParams(f.N, f.k, tlwe_alpha_min, tlwe_alpha_max);
    TGswParams* bkp = new_TGswParams(f.l, f.Bgbit, accum);
    return new_TFheGateBootstrappingParameterSet(f.ks_t, f.ks_basebit, in_out, bkp);
}

}  // namespace tfhe

static void destroy_cloud(const TFheGateBootstrappingCloudKeySet* c) {
    forget_device_keys(c);
    const int n = c->params->in_out_params->n;
    const int kpl = c->params->tgsw_params->kpl;
    LweBootstrappingKey* bk = const_cast<LweBootstrappingKey*>(c->bk);
    LweBootstrappingKeyFFT* bkfft = const_cast<LweBootstrappingKeyFFT*>(c->bkFFT);
    if (bkfft) {
        TGswSampleFFT* s = const_cast<TGswSampleFFT*>(bkfft->bkFFT);
        for (int i = 0; i < n; ++i) destroy_tgsw_fft(s + i, kpl);
        delete[] s;
        delete bkfft;
    }
    if (bk) {
        for (int i = 0; i < n; ++i) destroy_tgsw(bk->bk + i, kpl);
        delete[] bk->bk;
        delete_LweKeySwitchKey(bk->ks);
        delete bk;
    }
}

extern "C" void delete_gate_bootstrapping_secret_keyset(TFheGateBootstrappingSecretKeySet* keyset) {
    if (!keyset) return;
    destroy_cloud(&keyset->cloud);
    delete_LweKey(const_cast<LweKey*>(keyset->lwe_key));
    delete_TGswKey(const_cast<TGswKey*>(keyset->tgsw_key));
    const TFheGateBootstrappingParameterSet* owned = tfhe::take_owned_params(keyset);
    delete keyset;
    if (owned) delete_gate_bootstrapping_parameters(const_cast<TFheGateBootstrappingParameterSet*>(owned));
}

// A cloud key set made by keygen is embedded in its secret key set (as upstream does), so deleting
// only the cloud part releases its device keys and the rest goes with the secret key set.  A cloud
// key set read from a file stands alone and owns everything, its parameter set included.
extern "C" void delete_gate_bootstrapping_cloud_keyset(TFheGateBootstrappingCloudKeySet* keyset) {
    if (!keyset) return;
    forget_device_keys(keyset);
    const TFheGateBootstrappingParameterSet* owned = nullptr;
    if (tfhe::take_standalone_cloud(keyset, &owned)) {
        destroy_cloud(keyset);
        delete keyset;
        delete_gate_bootstrapping_parameters(const_cast<TFheGateBootstrappingParameterSet*>(owned));
    }
}

// ------------------------------------------------------------------------------------------
// flat export (used by the flat ABI, the tests and the benchmark)
// ------------------------------------------------------------------------------------------
namespace tfhe {

bool flat_params_of(const TFheGateBootstrappingParameterSet* p, tfhe_flat_params* out) {
    if (!p || !out) return false;
    const TGswParams* g = p->tgsw_params;
    *out = tfhe_flat_params{p->in_out_params->n, g->tlwe_params->N, g->tlwe_params->k, g->l, g->Bgbit,
                            p->ks_t, p->ks_basebit};
    return true;
}

void flatten_bk(const LweBootstrappingKey* bk, int32_t* out) {
    const int n = bk->in_out_params->n, N = bk->accum_params->N, k = bk->accum_params->k;
    const int kpl = bk->bk_params->kpl;
    size_t o = 0;
    for (int i = 0; i < n; ++i)
        for (int r = 0; r < kpl; ++r)
            for (int q = 0; q <= k; ++q) {
                memcpy(out + o, bk->bk[i].all_sample[r].a[q].coefsT, sizeof(int32_t) * N);
                o += N;
            }
}

void flatten_ksk(const LweKeySwitchKey* ks, int32_t* out) {
    const int n_out = ks->out_params->n;
    size_t o = 0;
    for (int i = 0; i < ks->n; ++i)
        for (int j = 0; j < ks->t; ++j)
            for (int h = 0; h < ks->base; ++h) {
                const LweSample* s = &ks->ks[i][j][h];
                memcpy(out + o, s->a, sizeof(int32_t) * n_out);
                out[o + n_out] = s->b;
                o += n_out + 1;
            }
}

}  // namespace tfhe

extern "C" int tfhe_keyset_export_flat(const TFheGateBootstrappingSecretKeySet* key, int32_t* lwe_key,
                                       int32_t* tlwe_key, int32_t* bk, int32_t* ksk) {
    if (!key) return TFHE_E_ARG;
    const int n = key->params->in_out_params->n;
    const TLweParams* tp = key->params->tgsw_params->tlwe_params;
    if (lwe_key) memcpy(lwe_key, key->lwe_key->key, sizeof(int32_t) * n);
    if (tlwe_key)
        for (int q = 0; q < tp->k; ++q) memcpy(tlwe_key + (size_t)q * tp->N, key->tgsw_key->tlwe_key.key[q].coefs,
                                               sizeof(int32_t) * tp->N);
    if (bk) flatten_bk(key->cloud.bk, bk);
    if (ksk) flatten_ksk(key->cloud.bk->ks, ksk);
    return TFHE_OK;
}

extern "C" int tfhe_cloud_export_flat(const TFheGateBootstrappingCloudKeySet* cloud, int32_t* bk, int32_t* ksk) {
    if (!cloud || !cloud->bk) return TFHE_E_ARG;
    if (bk) flatten_bk(cloud->bk, bk);
    if (ksk) flatten_ksk(cloud->bk->ks, ksk);
    return TFHE_OK;
}

extern "C" int tfhe_flat_params_of(const TFheGateBootstrappingParameterSet* params, tfhe_flat_params* out) {
    return flat_params_of(params, out) ? TFHE_OK : TFHE_E_ARG;
}

// ------------------------------------------------------------------------------------------
// gates (cloud.c:10-35 call shapes; alice.c:36 / verify.c:27 for encrypt/decrypt)
// ------------------------------------------------------------------------------------------
static const Torus32 kMU = (Torus32)(1u << 29);  // modSwitchToTorus32(1, 8), tb :267

extern "C" void bootsSymEncrypt(LweSample* result, int32_t message, const TFheGateBootstrappingSecretKeySet* key) {
    const Torus32 mu = message ? kMU : -kMU;
    lweSymEncrypt(result, mu, key->params->in_out_params->alpha_min, key->lwe_key);
}

extern "C" int32_t bootsSymDecrypt(const LweSample* sample, const TFheGateBootstrappingSecretKeySet* key) {
    return lwePhase(sample, key->lwe_key) > 0 ? 1 : 0;
}

extern "C" void bootsCONSTANT(LweSample* result, int32_t value, const TFheGateBootstrappingCloudKeySet* bk) {
    lweNoiselessTrivial(result, value ? kMU : -kMU, bk->params->in_out_params);
}
extern "C" void bootsNOT(LweSample* result, const LweSample* ca, const TFheGateBootstrappingCloudKeySet* bk) {
    lweNegate(result, ca, bk->params->in_out_params);
}
extern "C" void bootsCOPY(LweSample* result, const LweSample* ca, const TFheGateBootstrappingCloudKeySet* bk) {
    lweCopy(result, ca, bk->params->in_out_params);
}

static void run_gate(int32_t gate, LweSample* result, const LweSample* a, const LweSample* b, const LweSample* c,
                     const TFheGateBootstrappingCloudKeySet* bk) {
    const int rc = tfhe_gpu_boots_batch(gate, result, a, b, c, 1, bk);
    if (rc != TFHE_OK) fatal("boots gate", rc);
}

extern "C" void bootsNAND(LweSample* r, const LweSample* a, const LweSample* b, const TFheGateBootstrappingCloudKeySet* bk) { run_gate(TFHE_GATE_NAND, r, a, b, nullptr, bk); }
extern "C" void bootsAND(LweSample* r, const LweSample* a, const LweSample* b, const TFheGateBootstrappingCloudKeySet* bk) { run_gate(TFHE_GATE_AND, r, a, b, nullptr, bk); }
extern "C" void bootsOR(LweSample* r, const LweSample* a, const LweSample* b, const TFheGateBootstrappingCloudKeySet* bk) { run_gate(TFHE_GATE_OR, r, a, b, nullptr, bk); }
extern "C" void bootsXOR(LweSample* r, const LweSample* a, const LweSample* b, const TFheGateBootstrappingCloudKeySet* bk) { run_gate(TFHE_GATE_XOR, r, a, b, nullptr, bk); }
extern "C" void bootsXNOR(LweSample* r, const LweSample* a, const LweSample* b, const TFheGateBootstrappingCloudKeySet* bk) { run_gate(TFHE_GATE_XNOR, r, a, b, nullptr, bk); }
extern "C" void bootsNOR(LweSample* r, const LweSample* a, const LweSample* b, const TFheGateBootstrappingCloudKeySet* bk) { run_gate(TFHE_GATE_NOR, r, a, b, nullptr, bk); }
extern "C" void bootsANDNY(LweSample* r, const LweSample* a, const LweSample* b, const TFheGateBootstrappingCloudKeySet* bk) { run_gate(TFHE_GATE_ANDNY, r, a, b, nullptr, bk); }
extern "C" void bootsANDYN(LweSample* r, const LweSample* a, const LweSample* b, const TFheGateBootstrappingCloudKeySet* bk) { run_gate(TFHE_GATE_ANDYN, r, a, b, nullptr, bk); }
extern "C" void bootsORNY(LweSample* r, const LweSample* a, const LweSample* b, const TFheGateBootstrappingCloudKeySet* bk) { run_gate(TFHE_GATE_ORNY, r, a, b, nullptr, bk); }
extern "C" void bootsORYN(LweSample* r, const LweSample* a, const LweSample* b, const TFheGateBootstrappingCloudKeySet* bk) { run_gate(TFHE_GATE_ORYN, r, a, b, nullptr, bk); }
extern "C" void bootsMUX(LweSample* r, const LweSample* a, const LweSample* b, const LweSample* c,
                         const TFheGateBootstrappingCloudKeySet* bk) {
    run_gate(TFHE_GATE_MUX, r, a, b, c, bk);
}

extern "C" int tfhe_encrypt_bits_flat(const TFheGateBootstrappingSecretKeySet* key, const int32_t* bits,
                                      int32_t count, int32_t* out) {
    if (!key || count < 0 || (count > 0 && (!bits || !out))) return TFHE_E_ARG;
    const int n = key->params->in_out_params->n;
    LweSample* s = new_LweSample(key->params->in_out_params);
    for (int i = 0; i < count; ++i) {
        bootsSymEncrypt(s, bits[i], key);
        memcpy(out + (size_t)i * (n + 1), s->a, sizeof(int32_t) * n);
        out[(size_t)i * (n + 1) + n] = s->b;
    }
    delete_LweSample(s);
    return TFHE_OK;
}

extern "C" int tfhe_decrypt_bits_flat(const TFheGateBootstrappingSecretKeySet* key, const int32_t* samples,
                                      int32_t count, int32_t* bits) {
    if (!key || count < 0 || (count > 0 && (!bits || !samples))) return TFHE_E_ARG;
    const int n = key->params->in_out_params->n;
    const int32_t* s = key->lwe_key->key;
    for (int i = 0; i < count; ++i) {
        const int32_t* x = samples + (size_t)i * (n + 1);
        uint32_t axs = 0;
        for (int j = 0; j < n; ++j) axs += (uint32_t)x[j] * (uint32_t)s[j];
        bits[i] = ((int32_t)((uint32_t)x[n] - axs) > 0) ? 1 : 0;
    }
    return TFHE_OK;
}

// ------------------------------------------------------------------------------------------
// bootstrapping internals on upstream types
// ------------------------------------------------------------------------------------------
static void lwe_to_flat(const LweSample* s, int n, int32_t* out) {
    memcpy(out, s->a, sizeof(int32_t) * n);
    out[n] = s->b;
}
static void flat_to_lwe(const int32_t* in, int n, LweSample* s) {
    memcpy(s->a, in, sizeof(int32_t) * n);
    s->b = in[n];
    s->current_variance = 0.;
}

extern "C" void tfhe_bootstrap_FFT(LweSample* result, const LweBootstrappingKeyFFT* bk, Torus32 mu, const LweSample* x) {
    tfhe_gpu_key* key = device_key_for_lwebk(bk);
    if (!key) fatal("tfhe_bootstrap_FFT: key not created by this library", TFHE_E_ARG);
    const int n = bk->in_out_params->n;
    std::vector<int32_t> in(n + 1), out(n + 1);
    lwe_to_flat(x, n, in.data());
    const int rc = tfhe_gpu_bootstrap_batch(key, default_precision(), out.data(), in.data(), mu, 1);
    if (rc != TFHE_OK) fatal("tfhe_bootstrap_FFT", rc);
    flat_to_lwe(out.data(), n, result);
}

extern "C" void tfhe_bootstrap_woKS_FFT(LweSample* result, const LweBootstrappingKeyFFT* bk, Torus32 mu,
                                        const LweSample* x) {
    tfhe_gpu_key* key = device_key_for_lwebk(bk);
    if (!key) fatal("tfhe_bootstrap_woKS_FFT: key not created by this library", TFHE_E_ARG);
    const int n = bk->in_out_params->n, kN = bk->extract_params->n;
    std::vector<int32_t> in(n + 1), out(kN + 1);
    lwe_to_flat(x, n, in.data());
    const int rc = tfhe_gpu_bootstrap_woks_batch(key, default_precision(), out.data(), in.data(), mu, 1);
    if (rc != TFHE_OK) fatal("tfhe_bootstrap_woKS_FFT", rc);
    flat_to_lwe(out.data(), kN, result);
}

extern "C" void tfhe_blindRotate_FFT(TLweSample* accum, const TGswSampleFFT* bk, const int32_t* bara, int32_t n,
                                     const TGswParams* bk_params) {
    tfhe_gpu_key* key = device_key_for_bkfft(bk);
    if (!key) fatal("tfhe_blindRotate_FFT: key not created by this library", TFHE_E_ARG);
    const int N = bk_params->tlwe_params->N;
    tfhe_flat_params fp;
    key_params(key, &fp);
    const int32_t nkey = fp.n;
    if (n > nkey) fatal("tfhe_blindRotate_FFT: n exceeds the key length", TFHE_E_ARG);
    std::vector<int32_t> acc((size_t)2 * N), ba(nkey, 0);
    for (int q = 0; q < 2; ++q) memcpy(acc.data() + (size_t)q * N, accum->a[q].coefsT, sizeof(int32_t) * N);
    for (int i = 0; i < n; ++i) ba[i] = bara[i];
    const int rc = tfhe_gpu_blind_rotate_batch(key, default_precision(), acc.data(), ba.data(), 1);
    if (rc != TFHE_OK) fatal("tfhe_blindRotate_FFT", rc);
    for (int q = 0; q < 2; ++q) memcpy(accum->a[q].coefsT, acc.data() + (size_t)q * N, sizeof(int32_t) * N);
}

extern "C" void tLweExtractLweSample(LweSample* result, const TLweSample* x, const LweParams* params,
                                     const TLweParams* rparams) {
    const int N = rparams->N, k = rparams->k;
    (void)params;
    for (int q = 0; q < k; ++q) {
        result->a[q * N] = x->a[q].coefsT[0];
        for (int j = 1; j < N; ++j) result->a[q * N + j] = (int32_t)(0u - (uint32_t)x->a[q].coefsT[N - j]);
    }
    result->b = x->b->coefsT[0];
    result->current_variance = x->current_variance;
}

extern "C" void tfhe_blindRotateAndExtract_FFT(LweSample* result, const TorusPolynomial* v, const TGswSampleFFT* bk,
                                               int32_t barb, const int32_t* bara, int32_t n,
                                               const TGswParams* bk_params) {
    const TLweParams* accp = bk_params->tlwe_params;
    const int N = accp->N;
    TLweSample* acc = new_TLweSample(accp);
    // testvectbis = X^{2N - barb} v  (mulByXaiOrCopy.cpp:25-41), acc = (0, testvectbis)
    const int s = (2 * N - barb) % (2 * N);
    for (int j = 0; j < N; ++j) {
        const int e = ((j - s) % (2 * N) + 2 * N) % (2 * N);
        const uint32_t val = (uint32_t)v->coefsT[e % N];
        acc->b->coefsT[j] = (int32_t)(e < N ? val : 0u - val);
    }
    tfhe_blindRotate_FFT(acc, bk, bara, n, bk_params);
    tLweExtractLweSample(result, acc, &accp->extracted_lweparams, accp);
    delete_TLweSample(acc);
}

extern "C" void lweKeySwitch(LweSample* result, const LweKeySwitchKey* ks, const LweSample* sample) {
    tfhe_gpu_key* key = device_key_for_ksk(ks);
    if (!key) fatal("lweKeySwitch: key not created by this library", TFHE_E_ARG);
    const int kN = ks->n, n = ks->out_params->n;
    std::vector<int32_t> in(kN + 1), out(n + 1);
    lwe_to_flat(sample, kN, in.data());
    const int rc = tfhe_gpu_keyswitch_batch(key, out.data(), in.data(), 1);
    if (rc != TFHE_OK) fatal("lweKeySwitch", rc);
    flat_to_lwe(out.data(), n, result);
}
