// engine.h -- host-side interface of the gfx950 kernels (launchers live in engine.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tfhe {

enum Prec { PREC_FP64 = 0, PREC_SPLIT32 = 1 };

// Device parameter set (k = 1 only).
struct DevParams {
    int n, N, l, Bgbit, ks_t, ks_basebit;
};

// Per-transform-size twiddle tables in device memory (built once per N).
struct DevTwiddles {
    int N = 0;
    void* t1_fold64 = nullptr;   // [64][R1] C2<double>: w^{L(1+4k1)}  (twist folded, FP64 path)
    void* t2_64 = nullptr;       // [64][8]  C2<double>: W_64^{(L&7) k2}
    void* t1_plain32 = nullptr;  // [64][R1] C2<float>: W_M^{L k1}     (split path)
    void* t2_32 = nullptr;       // [64][8]  C2<float>
    void* t2dit32 = nullptr;     // [64][8]  C2<float>: W_64^{(L>>3) b}  (split path, pass-2 inputs)
    void* twist32 = nullptr;     // [64][R1] C2<float>:  w^{64a + L}      (split path, forward twist)
    void* untw64 = nullptr;      // [64][R1] C2<double>: w^{64a + L} / M  (split path, untwist)
    void* wa64 = nullptr;        // [R1]     C2<double>: w^{64a}
    // N = 1024 split path, forward transform of digits with its first pass on the matrix cores
    // (fft512_fwd_mfma): this lane's FP16 pieces of the real 16x16 pass-1 matrix, and the pass-2
    // input twiddles with the lane twist folded in
    void* fwd_a16 = nullptr;     // [64][16] FP16: this lane's entries of the two matrix pieces
    void* fwd_t2 = nullptr;      // [64][8] C2<float>: W_64^{(L>>3) b} w^{8b + (L&7)}
    // N = 2048, one transform over two wavefronts (wavefront v holds the frequencies of parity v):
    void* t1w64 = nullptr;       // [2][64][8] C2<double>: w^{L (1 + 4v + 8 k1)}
    void* t1w32 = nullptr;       // [2][64][8] C2<float>
    void* rw64 = nullptr;        // [2][8] C2<double>: w^{64 a (1 + 4v)}
    void* rw32 = nullptr;        // [2][8] C2<float>
    void* rws64 = nullptr;       // [2][8] C2<double>: w^{64 a (1 + 4v)} / 1024 (split path, own untwist)
    // N = 2048 split path, matrix-core form of each wavefront's forward transform
    // (fft512_fwd2w_mfma): this lane's entries of the two FP16 pieces of parity v's 16x32 matrix, and
    // the pass-1 twiddles in pass 2's layout (shared with the inverse, fft512_inv2w_t1p)
    void* w2_a16 = nullptr;      // [2][64][2][8] FP16: piece p of A_v[L & 15][8 (L >> 4) + i]
    void* w2_t1p32 = nullptr;    // [2][64][8] C2<float>: w^{(8b + (L&7)) (1 + 4v + 8 (L>>3))}
    // N = 1024, 4 bootstraps per wavefront (form 6): pass-2 twiddles [16][2][16] C2<float>
    void* q4tw = nullptr;
};

// One gate segment of a table-driven bootstrap launch (mode 3): jobs [s*seg_jobs, (s+1)*seg_jobs)
// read x = (0, cst) + alpha*a[i] + beta*b[i] for i = job - s*seg_jobs (b may be null).
struct SegDesc {
    const int32_t* a;
    const int32_t* b;
    int32_t alpha, beta, cst, pad;
};

// One LWE linear job (no bootstrap): dst[i] = mul * src[i] + (0, add_b) over `count` samples, or
// the trivial sample (0, add_b) when src is null.
struct LinDesc {
    const int32_t* src;
    int32_t* dst;
    int32_t mul, add_b;
};

// One launch of the fused bootstrap kernel.
//   mode 0: job g reads x = (0, cst[s]) + alpha[s]*in_a[s][g'] + beta[s]*in_b[s][g'] from segment s
//           (g' = g - seg_start[s]), modulus-switches it, builds the X^{-barb} test vector of mu,
//           blind-rotates and writes the extracted sample u_out[g][0..N].
//   mode 1: acc_io[g] (k+1 = 2 polys) is blind-rotated by bara_in[g][0..n) in place.
//   mode 2: acc_io[g] <- bk[row_set[g]] (.) acc_io[g] (one external product); raw_out, raw_hi optional.
//   mode 3: as mode 0, with the gate inputs of job g taken from segs[g / seg_jobs].
// form: 0 = the process-wide choice (set_forced_form, else the size rules);
//   1 = FP64 at N=1024: key-sharing workgroups (k_boot64w);
//   2 = one bootstrap per workgroup: k_boot64 / k_boot32 (N=1024), k_boot2w with one two-wavefront
//       bootstrap per workgroup (N=2048);
//   3 = latency form, one bootstrap per workgroup of several wavefronts (modes 0, 1, 3);
//   4, 5 = N=2048: k_boot2w with 2 or 4 bootstraps per workgroup;
//   6 = split FP32 at N=1024: k_boot32q, 4 bootstraps per wavefront with the products on the
//       matrix cores (modes 0, 1, 3; needs the key's bk32q).
// A form that does not apply to a launch's path (1 on the split path or at N=2048, 4 and 5 at
// N=1024) is treated as 0.
struct BootLaunch {
    int mode = 0;
    int form = 0;
    const SegDesc* segs = nullptr;
    int seg_jobs = 0;
    int count = 0;
    int nseg = 1;
    int seg_start[3] = {0, 0, 0};
    const int32_t* in_a[2] = {nullptr, nullptr};
    const int32_t* in_b[2] = {nullptr, nullptr};
    int32_t alpha[2] = {1, 1}, beta[2] = {0, 0}, cst[2] = {0, 0};
    int32_t mu = 0;
    int32_t* acc_io = nullptr;
    const int32_t* bara_in = nullptr;
    const int32_t* row_set = nullptr;
    double* raw_out = nullptr;
    double* raw_hi = nullptr;   // mode 2, split FP32: the hi product before rounding (units of 2^split)
    int32_t* u_out = nullptr;
    // fused keyswitch (split FP32, N = 1024, mode 0, one segment): job g writes the keyswitched
    // sample KS((0, ks_add_b) + extract(acc_g)) to ks_out[g] instead of the extracted sample
    int32_t* ks_out = nullptr;
    int32_t ks_add_b = 0;
};

struct DevKeyView {
    DevParams p;
    const void* bk64 = nullptr;      // FP64 FFT-domain key, lane order
    const void* bk32 = nullptr;      // split FP32 key {hi.re, hi.im, lo.re, lo.im}
    const void* bk32q = nullptr;     // the split key arranged for form 6 (N = 1024; may be null)
    int split_bits = 24;
    const int32_t* ksk = nullptr;    // [kN][t/2][15][NP] digit-pair rows
    int ksk_np = 0;                  // padded output length (multiple of 64)
};

// Fixed device scratch of the no-allocation gate entry point: kScratchElems int32 (16 MiB), one
// copy per device for the life of the process.  scratch_init resolves its address on the current
// device (call once per device at set-up); scratch_ptr returns it (null before scratch_init).
constexpr size_t kScratchElems = (size_t)4 << 20;
constexpr int kScratchMaxDevices = 64;
hipError_t scratch_init(int device);
int32_t* scratch_ptr(int device);

hipError_t build_twiddles(DevTwiddles& tw, int N);
void free_twiddles(DevTwiddles& tw);

bool params_supported(const DevParams& p);
// FP64 launch form when a launch names none: 0 one bootstrap per workgroup, 1 key-sharing
// workgroups, 2 by launch size
void set_boot64_shared(int mode);
// split-FP32 MAC at N = 1024: 0 plain FMA accumulation, 1 compensated (Dot2) accumulation
void set_split_mac(int mode);
// launch form for launches that name none (0 restores the size rules)
void set_forced_form(int form);
int forced_form();
// 1: split-FP32 gate batches at N = 1024 (not MUX) keyswitch inside the bootstrap kernel
void set_fused_keyswitch(int on);
int fused_keyswitch();
hipError_t launch_bootstrap(const DevKeyView& key, const DevTwiddles& tw, Prec prec, const BootLaunch& L,
                            hipStream_t st);
// KS((0, add_b) + u1 + u2) for each of count samples; u2 may be null.  Output g goes to out[g], or,
// with out_tab, to sample g mod out_seg of the device array out_tab[g / out_seg].
hipError_t launch_keyswitch(const DevKeyView& key, int32_t* out, const int32_t* u1, const int32_t* u2,
                            int32_t add_b, int count, hipStream_t st, int32_t* const* out_tab = nullptr,
                            int out_seg = 0);
// time-domain bk [n][2l][2][N] (device) -> FP64 and/or split FP32 FFT-domain keys
hipError_t launch_keyprep(const DevParams& p, const DevTwiddles& tw, const int32_t* d_bk, void* bk64, void* bk32,
                          int split_bits, hipStream_t st);
// KSK [kN][t][base][n+1] (device) -> [kN][t/2][15][NP]: row 4 d1 + d2 - 1 of pair g holds
// ks[2g][d1] + ks[2g+1][d2] (base 4 only)
size_t ksk_device_bytes(const DevParams& p, int np);
// split key bk32 -> the layout of launch form 6 (N = 1024, l = 3): n * 6 * 32 * 64 float4
hipError_t launch_keyprep_q4(const DevParams& p, const void* bk32, void* bk32q, hipStream_t st);
// transform test of form 6's 4-per-wavefront layout (N = 1024), semantics of launch_fft_test
hipError_t launch_fft_test_q4(const DevTwiddles& tw, int inverse, double* out, const void* in, int count,
                              hipStream_t st);
hipError_t launch_ksk_relayout(const DevParams& p, const int32_t* d_ksk_in, int32_t* d_ksk_out, int np,
                               hipStream_t st);
hipError_t launch_lwe_linear(const LinDesc* d_descs, int ndesc, int count, int n, hipStream_t st);
hipError_t launch_extract(int N, int32_t* u, const int32_t* acc, int count, hipStream_t st);
hipError_t launch_rotate_decompose(const DevParams& p, int32_t* rot, int32_t* digits, const int32_t* acc,
                                   const int32_t* bara, int count, hipStream_t st);
// inverse: 0 forward, 1 inverse, 2 the split path's forward of digit inputs with its first pass on
// the matrix cores (split FP32 only)
hipError_t launch_fft_test(const DevTwiddles& tw, int N, Prec prec, int inverse, double* out, const void* in,
                           int count, hipStream_t st);
// the split kernels' FP16 digit conversion (digits7_f16x2 / digits7_f16x2_one at bg = 7, variant 0
// / 1; digits8_f16x2 at bg = 8) of coefficient pairs x[2i], x[2i + 1]: out[i][lvl] = the two digits
hipError_t launch_digits_test(int bg, int variant, float* out, const int32_t* x, int count, hipStream_t st);
hipError_t launch_valu_probe(int fp64, void* out, int blocks, int iters, hipStream_t st);
hipError_t launch_decrypt(int32_t* bits, const int32_t* samples, const int32_t* key, int n, int count,
                          hipStream_t st);

}  // namespace tfhe
