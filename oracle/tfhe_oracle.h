/*
 * tfhe_oracle.h -- CPU oracle for the TFHE gate-bootstrapping path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it.
 * The product library (fpga-tfhe_amd) never links, loads or calls it.
 *
 * Every function restates the reference algorithm on flat integer arrays and
 * cites the reference file:line it follows (paths relative to the reference
 * repository root).  Where the reference delegates to the upstream `tfhe`
 * library (not vendored), the restatement follows the reference's call sites
 * and the published CGGI algorithm; such functions say so.
 *
 * Flat layouts (all int32, little endian, Torus32 = int32 mod 2^32):
 *   LWE sample       : [n+1]            a[0..n-1], b
 *   TLWE sample      : [(k+1)][N]        a[0..k-1], b (= a[k])
 *   TGSW sample      : [(k+1)*l][(k+1)][N], row r = bloc*l + level
 *   bootstrapping key: [n][(k+1)*l][(k+1)][N]   (time domain)
 *   keyswitch key    : [k*N][t][base][n+1]
 */
#ifndef TFHE_ORACLE_H
#define TFHE_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
    int32_t n;          /* LWE dimension                        */
    int32_t N;          /* TLWE ring degree                     */
    int32_t k;          /* TLWE mask polynomials                */
    int32_t l;          /* TGSW decomposition length            */
    int32_t Bgbit;      /* log2 of decomposition base           */
    int32_t ks_t;       /* keyswitch decomposition length       */
    int32_t ks_basebit; /* log2 of keyswitch base               */
    double  lwe_alpha;  /* LWE / keyswitch-key noise stdev      */
    double  bk_alpha;   /* bootstrapping-key noise stdev        */
} orc_params;

/* ---- torus numerics (kernel/src/numeric-functions.cpp) ---- */
void    orc_set_seed(const uint32_t* values, int32_t size);
int32_t orc_dtot32(double d);
double  orc_t32tod(int32_t x);
int32_t orc_gaussian32(int32_t message, double sigma);
int32_t orc_uniform_torus32(void);
int32_t orc_approx_phase(int32_t phase, int32_t Msize);
int32_t orc_modswitch_from_torus32(int32_t phase, int32_t Msize);
int32_t orc_modswitch_to_torus32(int32_t mu, int32_t Msize);
uint32_t orc_decomp_offset(int32_t l, int32_t Bgbit);

/* ---- keys (lwe-functions.cpp, lwe-keyswitch-functions.cpp, upstream CGGI) ---- */
void orc_keygen(const orc_params* p, const uint32_t* seed, int32_t seedlen,
                int32_t* lwe_key, int32_t* tlwe_key, int32_t* ksk, int32_t* bk);
void orc_lwe_encrypt(int32_t* out, int32_t message, double alpha, const int32_t* key, int32_t n);
int32_t orc_lwe_phase(const int32_t* sample, const int32_t* key, int32_t n);
void orc_boots_encrypt(const orc_params* p, int32_t* out, int32_t bit, const int32_t* lwe_key);
int32_t orc_boots_decrypt(const orc_params* p, const int32_t* sample, const int32_t* lwe_key);
void orc_tlwe_phase(const orc_params* p, int32_t* phase, const int32_t* sample, const int32_t* tlwe_key);

/* ---- polynomial primitives ---- */
void orc_poly_mul_xai(int32_t* res, int32_t a, const int32_t* src, int32_t N);
void orc_poly_mul_xai_minus_one(int32_t* res, int32_t a, const int32_t* src, int32_t N);
void orc_decompose(int32_t* res, const int32_t* src, int32_t N, int32_t l, int32_t Bgbit);
void orc_negacyclic_mul_i64(int64_t* res, const int32_t* ipoly, const int32_t* tpoly, int32_t N);
void orc_negacyclic_mul_naive(int32_t* res, const int32_t* ipoly, const int32_t* tpoly, int32_t N);

/* ---- FFT (folded negacyclic transform, FP64) ---- */
void orc_fft_forward(double* out_re_im, const int32_t* poly, int32_t N);
void orc_fft_inverse_raw(double* out, const double* in_re_im, int32_t N);
void orc_lagrange_half_reference(double* out_re_im, const int32_t* poly, int32_t N);
void orc_torus_from_lagrange_reference(int32_t* out, const double* in_re_im, int32_t N);

/* ---- external product / blind rotation ---- */
void orc_extern_product_exact(const orc_params* p, int32_t* acc, const int32_t* tgsw, int64_t* raw);
void* orc_fft_key_create(const orc_params* p, const int32_t* bk);
void orc_fft_key_free(void* h);
void orc_extern_product_fft(const orc_params* p, int32_t* acc, const void* fkey, int32_t row_set,
                            double* raw, int32_t truncate);
void orc_blind_rotate(const orc_params* p, int32_t* acc, const int32_t* bara, const int32_t* bk,
                      const void* fkey);
void orc_test_vector(const orc_params* p, int32_t* acc, int32_t barb, int32_t mu);
void orc_extract(const orc_params* p, int32_t* u, const int32_t* acc);
void orc_keyswitch(const orc_params* p, int32_t* res, const int32_t* ksk, const int32_t* u);
void orc_bootstrap_woks(const orc_params* p, int32_t* u, const int32_t* x, int32_t mu,
                        const int32_t* bk, const void* fkey);
void orc_bootstrap(const orc_params* p, int32_t* out, const int32_t* x, int32_t mu,
                   const int32_t* bk, const void* fkey, const int32_t* ksk);

/* ---- gates (upstream boots* semantics as used in test_bench/server/cloud.c) ---- */
enum {
    ORC_GATE_NAND = 0, ORC_GATE_AND, ORC_GATE_OR, ORC_GATE_XOR, ORC_GATE_XNOR,
    ORC_GATE_NOR, ORC_GATE_ANDNY, ORC_GATE_ANDYN, ORC_GATE_ORNY, ORC_GATE_ORYN,
    ORC_GATE_MUX
};
void orc_gate_combine(const orc_params* p, int32_t gate, int32_t* tmp, const int32_t* a, const int32_t* b);
void orc_gate(const orc_params* p, int32_t gate, int32_t* out, const int32_t* a, const int32_t* b,
              const int32_t* c, const int32_t* bk, const void* fkey, const int32_t* ksk);
double orc_gate_batch(const orc_params* p, int32_t gate, int32_t* out, const int32_t* a,
                      const int32_t* b, const int32_t* c, int32_t count, const void* fkey,
                      const int32_t* ksk, int32_t nthreads);

/* ---- planted mistakes, used only by the tests that prove the checks bite ---- */
void orc_set_fault(int32_t fault);

#ifdef __cplusplus
}
#endif
#endif
