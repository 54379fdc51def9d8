// engine.hip -- gfx950 kernels of the TFHE gate-bootstrapping path and their launchers.
//
// Kernels:
//   k_boot64 / k_boot32  fused bootstrap, one wavefront per bootstrap: gate pre-combination,
//                        modulus switch, test vector, n MuxRotate steps with the TLWE
//                        accumulator resident in LDS, sample extraction.  FP64 transform path /
//                        split-FP32 transform path.
//   k_boot64w            the FP64 bootstrap for large launches: workgroups of 8 bootstraps that
//                        share each key row through LDS.
//   k_boot64lat / k_boot32lat  latency forms for small launches: one bootstrap spread over a
//                        workgroup of 2l wavefronts, one per digit polynomial.
//   k_boot2w             N = 2048: each transform split over the two wavefronts of a bootstrap.
//   k_boot2wlat          N = 2048 latency form: one bootstrap per workgroup, a wavefront pair
//                        (the two spectrum halves) per digit level.
//   k_keyswitch          bit-exact LWE keyswitch, a tile of ciphertexts per workgroup so each
//                        keyswitch-key row is read once per tile (coalesced over coefficients);
//                        k_keyswitch_sliced spreads a few ciphertexts' input indices over
//                        workgroups instead.
//   k_keyprep            time-domain bootstrapping key -> FFT domain (FP64 or FP32 hi/lo).
//   k_ksk_relayout       keyswitch key -> [kN][t/2][15][NP] sums of digit-pair rows.
//   k_extract, k_rotdec, k_fft_test, k_decrypt  stage kernels for parity tests and audits.
#include <cmath>
#include <type_traits>
#include <cstring>
#include <vector>

#include "engine.h"
#include "tfhe_device.hpp"

namespace tfhe {

constexpr int kMaxLweN = 1024;   // largest LWE dimension the LDS prologue holds
constexpr int KS_PAIR_ROWS = 15;  // keyswitch key rows per digit pair: (d1, d2) in [0,4)^2 without (0, 0)

// Resident wavefronts per SIMD the bootstrap kernels are compiled for (register budget 512/W).
#ifndef TFHE_BOOT64_WAVES
#define TFHE_BOOT64_WAVES 1
#endif
#ifndef TFHE_BOOT32_WAVES
#define TFHE_BOOT32_WAVES 2
#endif
// Split path at N = 1024: 1 = the forward transforms of the digits take their first pass on the
// matrix cores (fft512_fwd_mfma), 0 = every pass on the vector ALU (fft512_dit)
#ifndef TFHE_FWD_MFMA
#define TFHE_FWD_MFMA 1
#endif
// the same in the split latency form (k_boot32lat)
#ifndef TFHE_LAT_FWD_MFMA
#define TFHE_LAT_FWD_MFMA 1
#endif
// Split path at N = 2048 (k_boot2w): 1 = each wavefront's forward transforms take the digit
// combination, register twist and first pass on the matrix cores (fft512_fwd2w_mfma), with the
// pass-1 twiddles in pass 2's layout shared by the inverse (fft512_inv2w_t1p); 0 = fft_fwd / fft_inv
#ifndef TFHE_W2_FWD_MFMA
#define TFHE_W2_FWD_MFMA 1
#endif
// N = 2048 split kernel with the matrix-core forward: the rotation reads each lane's coefficients
// in the matrix-core operand layout (w2_rotate_b), so no lane exchange precedes the matrix products.
// N = 2048 split latency form (k_boot2wlat): forward transforms with the matrix-core first pass (in
// the conventional layout, with the lane swaps) and digits by byte permutes; inverse transforms with
// the pass-1 table in pass 2's layout, as k_boot2w
#ifndef TFHE_W2LAT_FWD_MFMA
#define TFHE_W2LAT_FWD_MFMA 1
#endif
// 1: the split kernels' operand-layout rotations leave x + offset and the digits go to FP16 by byte
// permutes and one packed FP16 add (digits8_f16x2, digits7_f16x2); 0: signed bit-field extracts,
// int-to-float conversions and a packing conversion.
#ifndef DIGIT_PERM
#define DIGIT_PERM 1
#endif
#ifndef TFHE_W2_BLAYOUT
#define TFHE_W2_BLAYOUT TFHE_W2_FWD_MFMA
#endif

struct BootKArgs {
    BootLaunch L;
    const void* bk;
    const void* t1;
    const void* t2;
    const void* twist32;
    const void* untw64;
    const void* wa64;
    const void* t2w_rw64;   // N = 2048 two-wavefront register twists [2][8] (FP64)
    const void* t2w_rw32;   // and FP32
    int n;
    int split_bits;
    const int32_t* ksk;     // digit-pair keyswitch key (fused keyswitch form)
    const void* fwd_a16;    // N = 1024 split path: matrix-core pass-1 pieces (fft512_fwd_mfma)
    const void* fwd_t2;     // and its pass-2 table with the lane twist folded in
};

template <int N>
__device__ __forceinline__ constexpr int log2N() {
    return N == 1024 ? 10 : (N == 2048 ? 11 : (N == 512 ? 9 : 12));
}

// The gate inputs of bootstrap g in modes 0 and 3: the pre-combination x = cst + alpha a + beta b
// (b absent when xb is null), per input sample.
struct BootIn {
    const int32_t* xa;
    const int32_t* xb;
    uint32_t al, be;
    int32_t cst;
};
__device__ __forceinline__ BootIn boot_inputs(const BootKArgs& A, int g) {
    const int n = A.n;
    BootIn in;
    if (A.L.mode == 0) {
        // both entries are read and one is picked with a mask: a select of the two would be
        // folded into an indexed load, which copies the argument arrays to scratch
        const bool s1 = A.L.nseg > 1 && g >= A.L.seg_start[1];
        const uint32_t m = 0u - (uint32_t)s1;
        auto pick = [m](int32_t v0, int32_t v1) { return (int32_t)(((uint32_t)v0 & ~m) | ((uint32_t)v1 & m)); };
        auto pickp = [m](const int32_t* p0, const int32_t* p1) {
            const uint64_t mm = (uint64_t)(int64_t)(int32_t)m;
            return (const int32_t*)(((uint64_t)p0 & ~mm) | ((uint64_t)p1 & mm));
        };
        const int gl = g - pick(A.L.seg_start[0], A.L.seg_start[1]);
        const int32_t* ia = pickp(A.L.in_a[0], A.L.in_a[1]);
        const int32_t* ib = pickp(A.L.in_b[0], A.L.in_b[1]);
        in.xa = ia + (size_t)gl * (n + 1);
        in.xb = ib ? ib + (size_t)gl * (n + 1) : nullptr;
        in.al = (uint32_t)pick(A.L.alpha[0], A.L.alpha[1]);
        in.be = (uint32_t)pick(A.L.beta[0], A.L.beta[1]);
        in.cst = pick(A.L.cst[0], A.L.cst[1]);
    } else {
        const int s = g / A.L.seg_jobs;
        const int gl = g - s * A.L.seg_jobs;
        const SegDesc D = A.L.segs[s];
        in.xa = D.a + (size_t)gl * (n + 1);
        in.xb = D.b ? D.b + (size_t)gl * (n + 1) : nullptr;
        in.al = (uint32_t)D.alpha;
        in.be = (uint32_t)D.beta;
        in.cst = D.cst;
    }
    return in;
}

// Modulus switch to Z_2N: ((uint64(x) << 32) + interv/2) / interv with interv = 2^64 / 2N
// (tfhe_bootstrap_woKS_FFT.cpp:3-10) equals (x + 2^(SH-1)) >> SH on uint32, SH = 32 - log2(2N).
template <int N>
__device__ __forceinline__ int mod_switch_2n(uint32_t x) {
    constexpr int SH = 32 - (log2N<N>() + 1);
    return (int)((x + (1u << (SH - 1))) >> SH);
}

// Gate pre-combination + modulus switch + test vector (mode 0) or accumulator load (modes 1, 2).
// s_bara may be null for a kernel that recomputes the rotations at each step.
template <int N>
__device__ __forceinline__ void boot_prologue(const BootKArgs& A, int g, int lane, int32_t* s_acc,
                                              int16_t* s_bara) {
    const int n = A.n;
    if (A.L.mode == 0 || A.L.mode == 3) {
        const BootIn in = boot_inputs(A, g);
        if (s_bara) {
            for (int i = lane; i < n; i += 64) {
                uint32_t x = in.al * (uint32_t)in.xa[i];
                if (in.xb) x += in.be * (uint32_t)in.xb[i];
                s_bara[i] = (int16_t)mod_switch_2n<N>(x);
            }
        }
        uint32_t xbv = (uint32_t)in.cst + in.al * (uint32_t)in.xa[n];
        if (in.xb) xbv += in.be * (uint32_t)in.xb[n];
        const int barb = mod_switch_2n<N>(xbv);
        // b = X^{2N - barb} [mu, ..., mu] (mulByXaiOrCopy.cpp:25-41), a = 0 (tb :317)
        const int sft = (2 * N - barb) & (2 * N - 1);
        for (int j = lane; j < N; j += 64) {
            s_acc[j] = 0;
            const int e = (j - sft) & (2 * N - 1);
            s_acc[N + j] = (e < N) ? A.L.mu : -A.L.mu;
        }
    } else {
        const int32_t* ain = A.L.acc_io + (size_t)g * 2 * N;
        for (int j = lane; j < 2 * N; j += 64) s_acc[j] = ain[j];
        if (A.L.mode == 1 && s_bara)
            for (int i = lane; i < n; i += 64) s_bara[i] = (int16_t)A.L.bara_in[(size_t)g * n + i];
    }
}

// Sample extraction at index 0 (upstream tLweExtractLweSample, k = 1) or accumulator write-back.
template <int N>
__device__ __forceinline__ void boot_epilogue(const BootKArgs& A, int g, int lane, const int32_t* s_acc) {
    if (A.L.mode == 0 || A.L.mode == 3) {
        int32_t* u = A.L.u_out + (size_t)g * (N + 1);
        for (int j = lane; j < N; j += 64) u[j] = (j == 0) ? s_acc[0] : (int32_t)(0u - (uint32_t)s_acc[N - j]);
        if (lane == 0) u[N] = s_acc[N];
        if (A.L.acc_io) {
            int32_t* ao = A.L.acc_io + (size_t)g * 2 * N;
            for (int j = lane; j < 2 * N; j += 64) ao[j] = s_acc[j];
        }
    } else {
        int32_t* ao = A.L.acc_io + (size_t)g * 2 * N;
        for (int j = lane; j < 2 * N; j += 64) ao[j] = s_acc[j];
    }
}

// ---------------------------------------------------------------------------------------------
// FP64 path: one MuxRotate step (tfhe_MuxRotate_FFT.cpp:56-85) or one plain external product.
//   ROT:  acc <- acc + bk_i [.] ((X^a - 1) acc)
//   !ROT: acc <- bk_i [.] acc
// Decomposition and rotation are exact integer work; the external product runs the folded FP64
// transform with the twist folded into the pass-1 twiddles; the result is rounded to nearest and
// reduced mod 2^32 (the reference truncates, TorusPolynomial_fft.cpp:40).
//
// Key rows stream through an LDS buffer with the gfx950 direct global->LDS loads: the row that
// the next digit needs is requested as soon as the current row has been consumed, so its memory
// latency hides behind a full forward transform without holding any registers.
// ---------------------------------------------------------------------------------------------
typedef __attribute__((address_space(3))) void lds_void;

// all lanes request row `row` of the key row set `kr` (2 polynomials of M complex) into kbuf;
// lane L of the request for (poly, r) lands at element (poly*R1 + r)*64 + L
template <int M>
__device__ __forceinline__ void key_row_to_lds(C2<double>* kbuf, const C2<double>* __restrict__ kr, int row,
                                               int lane) {
    constexpr int R1 = M / 64;
    const C2<double>* src = kr + (size_t)row * 2 * M + lane;
    // the previous row's reads from kbuf must have returned before it is overwritten
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
#pragma unroll
    for (int pr = 0; pr < 2 * R1; ++pr)
        __builtin_amdgcn_global_load_lds((const void*)(src + pr * 64), (lds_void*)(kbuf + pr * 64), 16, 0, 0);
}

// KEY_LDS: key rows come through the LDS buffer kbuf (prefetched one row ahead); otherwise each
// lane loads its key values from memory right after the forward transform.  HALF: the transform
// exchanges go through a half-width buffer (real, then imaginary parts).
template <int N, int L, int BG, bool ROT, bool KEY_LDS = true, bool HALF = false>
__device__ __forceinline__ void step64(int32_t* s_acc, void* s_buf, C2<double>* kbuf, int a,
                                       const C2<double>* __restrict__ kr, const C2<double>* __restrict__ kr_next,
                                       const LaneTw<double, N / 128>& tw, const C2<double>* wa, int lane,
                                       double* raw) {
    constexpr int M = N / 2, R1 = M / 64, NC = 2 * R1;
    constexpr uint32_t OFF = decomp_offset<L, BG>();
    C2<double> o0[R1], o1[R1];
#pragma unroll
    for (int r = 0; r < R1; ++r) { o0[r] = C2<double>{0.0, 0.0}; o1[r] = C2<double>{0.0, 0.0}; }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        uint32_t u[NC];
        const uint32_t base4 = rot_base4(lane, a);
#pragma unroll
        for (int r = 0; r < NC; ++r) {
            const int j = lane + 64 * r;
            u[r] = (ROT ? rotdiff_e4<N>(s_acc + q * N, base4 + 256u * r, j, OFF) : (uint32_t)s_acc[q * N + j] + OFF) ^ OFF;
        }
#pragma unroll
        for (int lvl = 0; lvl < L; ++lvl) {
            const int row = q * L + lvl;
            C2<double> v[R1];
#pragma unroll
            for (int aa = 0; aa < R1; ++aa) {
                v[aa] = C2<double>{(double)digit_sx<BG>(u[aa], lvl), (double)digit_sx<BG>(u[aa + R1], lvl)};
                if (aa) v[aa] = cmul(v[aa], wa[aa]);
            }
            fft_fwd<R1, double, HALF>(v, tw, s_buf, lane);
            if constexpr (KEY_LDS) {
#pragma unroll
                for (int r = 0; r < R1; ++r) cmac(o0[r], v[r], kbuf[r * 64 + lane]);
#pragma unroll
                for (int r = 0; r < R1; ++r) cmac(o1[r], v[r], kbuf[(R1 + r) * 64 + lane]);
                if (row + 1 < 2 * L) key_row_to_lds<M>(kbuf, kr, row + 1, lane);
                else key_row_to_lds<M>(kbuf, kr_next, 0, lane);
            } else {
                const C2<double>* k = kr + (size_t)row * 2 * M + lane;
#pragma unroll
                for (int r = 0; r < R1; ++r) cmac(o0[r], v[r], k[r * 64]);
#pragma unroll
                for (int r = 0; r < R1; ++r) cmac(o1[r], v[r], k[(R1 + r) * 64]);
            }
        }
    }
    fft_inv<R1, double, HALF>(o0, tw, s_buf, lane);
    fft_inv<R1, double, HALF>(o1, tw, s_buf, lane);
    const double sc = 1.0 / (double)M;
#pragma unroll
    for (int aa = 0; aa < R1; ++aa) {
        const int j = lane + 64 * aa;
        C2<double> z0 = aa ? cmulc(o0[aa], wa[aa]) : o0[aa];
        C2<double> z1 = aa ? cmulc(o1[aa], wa[aa]) : o1[aa];
        z0.x *= sc; z0.y *= sc; z1.x *= sc; z1.y *= sc;
        if (raw) { raw[j] = z0.x; raw[j + M] = z0.y; raw[N + j] = z1.x; raw[N + j + M] = z1.y; }
        const uint32_t r00 = round_mod32(z0.x), r01 = round_mod32(z0.y);
        const uint32_t r10 = round_mod32(z1.x), r11 = round_mod32(z1.y);
        if (ROT) {
            s_acc[j] = (int32_t)((uint32_t)s_acc[j] + r00);
            s_acc[j + M] = (int32_t)((uint32_t)s_acc[j + M] + r01);
            s_acc[N + j] = (int32_t)((uint32_t)s_acc[N + j] + r10);
            s_acc[N + j + M] = (int32_t)((uint32_t)s_acc[N + j + M] + r11);
        } else {
            s_acc[j] = (int32_t)r00; s_acc[j + M] = (int32_t)r01;
            s_acc[N + j] = (int32_t)r10; s_acc[N + j + M] = (int32_t)r11;
        }
    }
    wave_sync();
}

template <int N, int L, int BG>
__global__ void __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(TFHE_BOOT64_WAVES, TFHE_BOOT64_WAVES))) k_boot64(BootKArgs A) {
    constexpr int M = N / 2, R1 = M / 64, KPL = 2 * L;
    __shared__ int32_t s_acc[2 * N];
    __shared__ C2<double> s_buf[xbuf_elems<M>()];
    __shared__ C2<double> s_key[2 * M];
    __shared__ int16_t s_bara[kMaxLweN];
    const int lane = threadIdx.x;
    const int g = blockIdx.x;
    boot_prologue<N>(A, g, lane, s_acc, s_bara);
    wave_sync();

    LaneTw<double, R1> tw;
    load_lane_tw(tw, (const C2<double>*)A.t1, (const C2<double>*)A.t2, lane);
    C2<double> wa[R1];
#pragma unroll
    for (int r = 0; r < R1; ++r) wa[r] = ((const C2<double>*)A.wa64)[r];
    const C2<double>* key = (const C2<double>*)A.bk;
    constexpr size_t ROWSET = (size_t)KPL * 2 * M;

    if (A.L.mode == 2) {
        double* raw = A.L.raw_out ? A.L.raw_out + (size_t)g * 2 * N : nullptr;
        const C2<double>* kr = key + (size_t)A.L.row_set[g] * ROWSET;
        key_row_to_lds<M>(s_key, kr, 0, lane);
        step64<N, L, BG, false>(s_acc, s_buf, s_key, 0, kr, kr, tw, wa, lane, raw);
    } else {
        const int n = A.n;
        if (n > 0) key_row_to_lds<M>(s_key, key, 0, lane);
        for (int i = 0; i < n; ++i) {
            const int a = __builtin_amdgcn_readfirstlane((int)s_bara[i]);
            // next key row set; the last step re-requests its own (valid) rows, which keeps the
            // prefetch unconditional
            const C2<double>* knext = key + (size_t)(i + 1 < n ? i + 1 : i) * ROWSET;
            if (a == 0) {
                // zero rotation: the step is skipped (tfhe_blindRotate_FFT.cpp:114); the buffer must
                // then hold row 0 of the next key instead (rare: one step in 2N)
                __builtin_amdgcn_s_waitcnt(0);
                key_row_to_lds<M>(s_key, knext, 0, lane);
                continue;
            }
            step64<N, L, BG, true>(s_acc, s_buf, s_key, a, key + (size_t)i * ROWSET, knext, tw, wa, lane, nullptr);
        }
    }
    // the last request (a harmless re-read of a valid row) must land before the LDS is released
    __builtin_amdgcn_s_waitcnt(0);
    boot_epilogue<N>(A, g, lane, s_acc);
}

// ---------------------------------------------------------------------------------------------
// The matrix-core operand layout of the split kernels' digits.  Register r of lane (n0, q) =
// (lane & 15, lane >> 4) holds coefficient bl_lane(lane) + bl_off(r): point j' = 16 g + n0 +
// 64 (2 q + t), rr = r & 7 = 2 g + t, of the 512-coefficient block r >> 3 (j' + 512 (r >> 3)).  The
// FP16 digit pairs of registers (rr, rr + 8) (N = 1024) or (rr, rr + 16) and (rr + 8, rr + 24)
// (N = 2048) are then the B operand of the first-pass matrix products as they stand, and no lane
// exchange precedes them.
// ---------------------------------------------------------------------------------------------
__host__ __device__ constexpr int bl_off(int r) { return 16 * ((r & 7) >> 1) + 64 * (r & 1) + 512 * (r >> 3); }
__device__ __forceinline__ int bl_lane(int lane) { return (lane & 15) + 128 * (lane >> 4); }

// ---------------------------------------------------------------------------------------------
// Rotation at N = 1024 from an accumulator with mirrored ends, as w2_rotate does at N = 2048: the
// MW = 64 (G - 1) words below p0 repeat p0's last MW coefficients and the MW past p1 repeat p1's
// first.  Registers r = G g + rr (r < 16) read p[(b_g + 64 rr) mod N], b_g = (d + 64 G g) mod N,
// d = lane - a: p1 directly (up to N + MW - 1), p0 from c_g = (b_g + MW) mod N at c_g + 64 rr - MW
// in [-MW, N), so every base register is non-negative.  The sign of register r is bit r of
// rotr32(0xFFFF0000, E), E = (d mod 2N) >> 6.  u[r] = ((X^a - 1) p_q + offset) ^ offset.  Used by
// the split latency form (G = 8) and the FP64 key-sharing form (G = 4, what its LDS has room for).
// ---------------------------------------------------------------------------------------------
template <int G>
__host__ __device__ constexpr int lat_mirror() { return 64 * (G - 1); }
template <int G>
__device__ __forceinline__ void lat_mirror_fill(int32_t* s_acc, int lane) {
    constexpr int N = 1024, MW = lat_mirror<G>();
    static_assert((N - MW) % 64 == 0, "mirror sources in the copying lane");
    // lane L copies coefficients it wrote itself in the prologue, so no fence is needed
    for (int k = lane; k < MW; k += 64) {
        s_acc[k - MW] = s_acc[N - MW + k];
        s_acc[2 * N + k] = s_acc[N + k];
    }
}
template <int G, uint32_t OFF, bool XOFF = true>
__device__ __forceinline__ void lat_rotate(uint32_t* u, const int32_t* s_acc, int q, int a, int lane) {
    constexpr int N = 1024, NC = 16, MW = lat_mirror<G>(), NG = NC / G;
    const int32_t* pq = s_acc + q * N;
    const int d = lane - a;
    const uint32_t sgn = __builtin_amdgcn_alignbit(0xFFFF0000u, 0xFFFF0000u, (uint32_t)(d >> 6) & 31u);
    const int sh = q ? 0 : MW;
    int base[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g) base[g] = (d + 64 * G * g + sh) & (N - 1);
#pragma unroll
    for (int r = 0; r < NC; ++r) {
        const uint32_t pj = (uint32_t)pq[lane + 64 * r];
        const uint32_t sv = (uint32_t)pq[base[r / G] + 64 * (r % G) - sh];
        const uint32_t m = (uint32_t)((int32_t)(sgn << (31 - r)) >> 31);
        uint32_t c = OFF - pj;
        asm("" : "+v"(c));
        u[r] = ((sv + m) ^ m) + c;
        if constexpr (XOFF) u[r] ^= OFF;
    }
}
// lat_rotate for the operand layout (bl_off, r < 16): register r reads p at e = (d + bl_off(r))
// mod 2N, d = bl_lane(lane) - a, negated when bit 10 of d + bl_off(r) is set.  Each half r >> 3
// reads from one base at offsets bl_off(r & 7) <= 112, so 128-word mirrors (lat_mirror<3>) cover
// every read.  The reads of one register put the four rows of 16 lanes 128 words apart.
template <int MW, uint32_t OFF, bool XOFF = true>
__device__ __forceinline__ void lat_rotate_b(uint32_t* u, const int32_t* s_acc, int q, int a, int lane) {
    constexpr int N = 1024, NC = 16;
    static_assert(MW >= 112, "mirrors cover the largest in-half offset");
    const int32_t* pq = s_acc + q * N;
    const int j0 = bl_lane(lane);
    const int d = j0 - a;
    const uint32_t x = (uint32_t)d << 21;   // bit 10 of d + offset -> bit 31
    const int sh = q ? 0 : MW;
    int base[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) base[k] = (d + 512 * k + sh) & (N - 1);
#pragma unroll
    for (int r = 0; r < NC; ++r) {
        const uint32_t pj = (uint32_t)pq[j0 + bl_off(r)];
        const uint32_t sv = (uint32_t)pq[base[r >> 3] + bl_off(r & 7) - sh];
        const uint32_t m = (uint32_t)((int32_t)(x + ((uint32_t)bl_off(r) << 21)) >> 31);
        uint32_t c = OFF - pj;
        asm("" : "+v"(c));
        u[r] = ((sv + m) ^ m) + c;
        if constexpr (XOFF) u[r] ^= OFF;
    }
}
// After coefficients j and j + N/2 of p_q (j = lane + 64 aa, aa < 8, the index within p_q) have been
// updated to lo and hi: p0's coefficients from N - MW up are hi values with aa >= 8 - MW / 64, p1's
// below MW are lo values with aa < MW / 64.
template <int G>
__device__ __forceinline__ void lat_mirror_update(int32_t* s_acc, int q, int aa, int j, int32_t lo, int32_t hi) {
    constexpr int N = 1024, MW = lat_mirror<G>();
    if (q == 0 && aa >= 8 - MW / 64) s_acc[j + N / 2 - N] = hi;
    if (q == 1 && aa < MW / 64) s_acc[2 * N + j] = lo;
}

// ---------------------------------------------------------------------------------------------
// FP64 path, latency form: one bootstrap per workgroup of 2l wavefronts.  In each step wavefront
// w = (q, lvl) rotates polynomial q of the accumulator, transforms its level-lvl digits and
// multiplies them by key row w, leaving the two partial products in LDS.  After a barrier,
// wavefronts 0 and 1 each sum the 2l partials of one output polynomial in a fixed order, invert it
// and add it into the accumulator.  A step's critical path is one forward and one inverse transform
// instead of eight transforms in sequence.  Each wavefront loads its key row for the next step
// into registers as soon as its products are written.  For launches too small to fill the GPU
// (single gates); its outputs equal the other FP64 forms' bit for bit, since each is the exact
// integer that the final rounding recovers.
// ---------------------------------------------------------------------------------------------
template <int N, int L, int BG>
__global__ void __launch_bounds__(64 * 2 * L) k_boot64lat(BootKArgs A) {
    constexpr int M = N / 2, R1 = M / 64, NC = 2 * R1, KPL = 2 * L;
    constexpr uint32_t OFF = decomp_offset<L, BG>();
    constexpr size_t ROW = (size_t)2 * M, ROWSET = (size_t)KPL * ROW;
    __shared__ int32_t s_acc[2 * N];
    __shared__ int16_t s_bara[kMaxLweN];
    __shared__ double s_x[KPL][xbuf_elems<M>()];   // half-width exchange buffer of each wavefront
    __shared__ C2<double> s_part[KPL][2][M];        // partial products [key row][output polynomial]
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int g = blockIdx.x;
    if (w == 0) boot_prologue<N>(A, g, lane, s_acc, s_bara);
    __syncthreads();

    LaneTw<double, R1> tw;
    load_lane_tw(tw, (const C2<double>*)A.t1, (const C2<double>*)A.t2, lane);
    C2<double> wa[R1];
#pragma unroll
    for (int r = 0; r < R1; ++r) wa[r] = ((const C2<double>*)A.wa64)[r];
    const int q = w / L, lvl = w % L;
    const C2<double>* krow = (const C2<double>*)A.bk + (size_t)w * ROW + lane;
    C2<double> k0[R1], k1[R1];
    auto load_row = [&](int i) {
        const C2<double>* p = krow + (size_t)i * ROWSET;
#pragma unroll
        for (int r = 0; r < R1; ++r) { k0[r] = p[r * 64]; k1[r] = p[M + r * 64]; }
    };
    const int n = A.n;
    int loaded = -1;
    for (int i = 0; i < n; ++i) {
        // the same value for every wavefront: the whole workgroup skips a zero rotation
        // (tfhe_blindRotate_FFT.cpp:114)
        const int a = __builtin_amdgcn_readfirstlane((int)s_bara[i]);
        if (a == 0) continue;
        if (loaded != i) load_row(i);
        {
            // the byte-offset rotation: the mirrored one (lat_rotate) spills registers in this kernel
            uint32_t u[NC];
            const uint32_t base4 = rot_base4(lane, a);
#pragma unroll
            for (int r = 0; r < NC; ++r)
                u[r] = rotdiff_e4<N>(s_acc + q * N, base4 + 256u * r, lane + 64 * r, OFF) ^ OFF;
            C2<double> v[R1];
#pragma unroll
            for (int aa = 0; aa < R1; ++aa) {
                v[aa] = C2<double>{(double)digit_sx<BG>(u[aa], lvl), (double)digit_sx<BG>(u[aa + R1], lvl)};
                if (aa) v[aa] = cmul(v[aa], wa[aa]);
            }
            fft_fwd<R1, double, true>(v, tw, s_x[w], lane);
#pragma unroll
            for (int r = 0; r < R1; ++r) {
                s_part[w][0][r * 64 + lane] = cmul(v[r], k0[r]);
                s_part[w][1][r * 64 + lane] = cmul(v[r], k1[r]);
            }
        }
        if (i + 1 < n) load_row(i + 1);
        loaded = i + 1;
        // the partials are complete, and every wavefront has read the accumulator for its rotation
        __syncthreads();
        if (w < 2) {
            C2<double> o[R1];
#pragma unroll
            for (int r = 0; r < R1; ++r) {
                o[r] = s_part[0][w][r * 64 + lane];
#pragma unroll
                for (int t = 1; t < KPL; ++t) o[r] = o[r] + s_part[t][w][r * 64 + lane];
            }
            fft_inv<R1, double, true>(o, tw, s_x[w], lane);
            const double sc = 1.0 / (double)M;
            int32_t* acc = s_acc + w * N;
#pragma unroll
            for (int aa = 0; aa < R1; ++aa) {
                const int j = lane + 64 * aa;
                const C2<double> z = aa ? cmulc(o[aa], wa[aa]) : o[aa];
                acc[j] = (int32_t)((uint32_t)acc[j] + round_mod32(z.x * sc));
                acc[j + M] = (int32_t)((uint32_t)acc[j + M] + round_mod32(z.y * sc));
            }
        }
        // the accumulator is updated and the partial buffers are free
        __syncthreads();
    }
    if (w == 0) boot_epilogue<N>(A, g, lane, s_acc);
}

// ---------------------------------------------------------------------------------------------
// FP64 path, key-sharing form.  A workgroup of B64W_GATES wavefronts runs that many bootstraps in
// lockstep over the key: every key row is brought into LDS once per workgroup, each wavefront
// loading its share with direct global->LDS loads, instead of once per bootstrap.  Per-bootstrap
// LDS (accumulator, half-width exchange buffer, bara) is small enough for two wavefronts per SIMD,
// which is what hides the transform's LDS latency.  Rows rotate through three buffers: row t+2 is
// requested right after the barrier of row t, when every wavefront has finished with row t-1, and
// has two transforms' time to land.
// ---------------------------------------------------------------------------------------------
constexpr int B64W_GATES = 8;
constexpr int B64W_MAXN = 640;   // largest LWE dimension this form is launched for
// the split kernel reads its digits in the matrix-core operand layout (lat_rotate_b) when its
// forward transforms take the matrix-core first pass
// 1: the split latency form reads its digits in the operand layout (lat_rotate_b) for its
// matrix-core forward; 0 (default): lat_rotate and the lane swaps.  Off because a lone bootstrap's
// transforms run back to back, so the operand layout's bank-conflicted reads sit on its critical
// path: 2.95 against 2.85 ms per single gate.
#ifndef LAT32_BLAYOUT
#define LAT32_BLAYOUT 0
#endif
#ifndef K32_BLAYOUT
#define K32_BLAYOUT (TFHE_FWD_MFMA && !FWD_MFMA_K32)
#endif
#ifndef K32_MIRROR_G
#define K32_MIRROR_G (K32_BLAYOUT ? 3 : 8)   // mirror group size of the split kernel's accumulator
#endif
#ifndef B64W_MIRROR_G
#define B64W_MIRROR_G 4   // mirror group size of the accumulators (lat_rotate)
#endif

template <int M, int W>
__device__ __forceinline__ void key_row_share(C2<double>* kbuf, const C2<double>* __restrict__ src, int wave,
                                              int lane) {
    constexpr int R1 = M / 64, PER = 2 * R1 / W;
    static_assert((2 * R1) % W == 0, "a key row splits evenly over the wavefronts");
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int pr = wave * PER + p;
        __builtin_amdgcn_global_load_lds((const void*)(src + pr * 64 + lane), (lds_void*)(kbuf + pr * 64), 16, 0, 0);
    }
}

// Workgroup execution barrier for the shared key rows.  The rows arrive through direct
// global->LDS loads, whose completion each wavefront establishes with its own vmcnt wait before
// the barrier; LDS is coherent across the CU, so no cache maintenance is needed.  The
// wavefront-scope fences only stop the compiler from moving LDS accesses across the barrier (a
// workgroup-scope fence would also drain every outstanding load, including the prefetch of the
// row after next).
__device__ __forceinline__ void rows_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// s_waitcnt with only vmcnt set (gfx9 encoding: vmcnt[3:0], expcnt[6:4], lgkmcnt[11:8], vmcnt[15:14])
template <int VM>
__device__ __forceinline__ void wait_vmcnt() {
    static_assert(VM >= 0 && VM < 16, "vmcnt");
    __builtin_amdgcn_s_waitcnt(0x0F70 | VM);
}

template <int N, int L, int BG>
__global__ void __launch_bounds__(64 * B64W_GATES)
__attribute__((amdgpu_waves_per_eu(2, 2))) k_boot64w(BootKArgs A) {
    static_assert(L == 3, "buffer rotation assumes 3 levels: row t of step i lands in buffer t mod 3 = lvl");
    constexpr int M = N / 2, R1 = M / 64, NC = 2 * R1, KPL = 2 * L, W = B64W_GATES;
    constexpr int PER = 2 * R1 / W;
    constexpr uint32_t OFF = decomp_offset<L, BG>();
    constexpr size_t ROW = (size_t)2 * M, ROWSET = (size_t)KPL * ROW;
    // accumulators with mirrored ends (lat_rotate); the modulus-switched rotations are not kept in
    // LDS but recomputed from the inputs 64 steps at a time (one per lane, read with v_readlane)
    constexpr int G = B64W_MIRROR_G, MW = lat_mirror<G>();
    __shared__ int32_t s_acc_all[W][2 * N + 2 * MW];
    __shared__ double s_x_all[W][xbuf_elems<M>()];
    __shared__ C2<double> s_k0[2 * M];
    __shared__ C2<double> s_k1[2 * M];
    __shared__ C2<double> s_k2[2 * M];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int g = blockIdx.x * W + wave;
    const bool active = g < A.L.count;
    int32_t* s_acc = s_acc_all[wave] + MW;
    double* s_x = s_x_all[wave];
    if (active) {
        boot_prologue<N>(A, g, lane, s_acc, nullptr);
        lat_mirror_fill<G>(s_acc, lane);
    }
    wave_sync();
    const int n = A.n;
    const BootIn in = (active && A.L.mode != 1) ? boot_inputs(A, g) : BootIn{nullptr, nullptr, 0u, 0u, 0};
    // rotation of step i (0 past the last step, or for an idle wavefront)
    auto bara_at = [&](int i) -> int {
        if (!active || i >= n) return 0;
        if (A.L.mode == 1) return A.L.bara_in[(size_t)g * n + i];
        uint32_t x = in.al * (uint32_t)in.xa[i];
        if (in.xb) x += in.be * (uint32_t)in.xb[i];
        return mod_switch_2n<N>(x);
    };
    int bcur = bara_at(lane), bnext = bara_at(64 + lane);   // steps 0-63 and 64-127

    LaneTw<double, R1> tw;
    load_lane_tw(tw, (const C2<double>*)A.t1, (const C2<double>*)A.t2, lane);
    C2<double> wa[R1];
#pragma unroll
    for (int r = 0; r < R1; ++r) wa[r] = ((const C2<double>*)A.wa64)[r];
    const C2<double>* key = (const C2<double>*)A.bk;
    C2<double>* const kb[3] = {s_k0, s_k1, s_k2};

    key_row_share<M, W>(s_k0, key, wave, lane);
    key_row_share<M, W>(s_k1, key + ROW, wave, lane);
    for (int i = 0; i < n; ++i) {
        if ((i & 63) == 0 && i > 0) {
            bcur = bnext;
            bnext = bara_at(i + 64 + lane);
        }
        const int a = __builtin_amdgcn_readlane(bcur, i & 63);
        // a zero rotation skips the step (tfhe_blindRotate_FFT.cpp:114); an idle wavefront still
        // takes part in the barriers and in loading the shared rows
        const bool work = a != 0;
        C2<double> o0[R1], o1[R1];
#pragma unroll
        for (int r = 0; r < R1; ++r) { o0[r] = C2<double>{0.0, 0.0}; o1[r] = C2<double>{0.0, 0.0}; }
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            uint32_t u[NC];
            if (work) {
                lat_rotate<G, OFF>(u, s_acc, q, a, lane);
            }
#pragma unroll
            for (int lvl = 0; lvl < L; ++lvl) {
                const int row = q * L + lvl;
                C2<double> v[R1];
                if (work) {
#pragma unroll
                    for (int aa = 0; aa < R1; ++aa) {
                        v[aa] = C2<double>{(double)digit_sx<BG>(u[aa], lvl), (double)digit_sx<BG>(u[aa + R1], lvl)};
                        if (aa) v[aa] = cmul(v[aa], wa[aa]);
                    }
                    fft_fwd<R1, double, true>(v, tw, s_x, lane);
                }
                // this wavefront's share of row t has landed (only row t+1's loads may remain),
                // and after the barrier every share has, and row t-1 is free
                wait_vmcnt<PER>();
                rows_barrier();
                {
                    const int r2 = row + 2;
                    const C2<double>* src = (r2 < KPL) ? key + (size_t)i * ROWSET + (size_t)r2 * ROW
                                                       : key + (size_t)(i + 1 < n ? i + 1 : i) * ROWSET +
                                                             (size_t)(r2 - KPL) * ROW;
                    key_row_share<M, W>(kb[(lvl + 2) % 3], src, wave, lane);
                }
                if (work) {
                    const C2<double>* k = kb[lvl];
#pragma unroll
                    for (int r = 0; r < R1; ++r) cmac(o0[r], v[r], k[r * 64 + lane]);
#pragma unroll
                    for (int r = 0; r < R1; ++r) cmac(o1[r], v[r], k[(R1 + r) * 64 + lane]);
                }
            }
        }
        if (work) {
            fft_inv<R1, double, true>(o0, tw, s_x, lane);
            fft_inv<R1, double, true>(o1, tw, s_x, lane);
            const double sc = 1.0 / (double)M;
#pragma unroll
            for (int aa = 0; aa < R1; ++aa) {
                const int j = lane + 64 * aa;
                C2<double> z0 = aa ? cmulc(o0[aa], wa[aa]) : o0[aa];
                C2<double> z1 = aa ? cmulc(o1[aa], wa[aa]) : o1[aa];
                const int32_t a0 = (int32_t)((uint32_t)s_acc[j] + round_mod32(z0.x * sc));
                const int32_t a1 = (int32_t)((uint32_t)s_acc[j + M] + round_mod32(z0.y * sc));
                const int32_t b0 = (int32_t)((uint32_t)s_acc[N + j] + round_mod32(z1.x * sc));
                const int32_t b1 = (int32_t)((uint32_t)s_acc[N + j + M] + round_mod32(z1.y * sc));
                s_acc[j] = a0;
                s_acc[j + M] = a1;
                s_acc[N + j] = b0;
                s_acc[N + j + M] = b1;
                lat_mirror_update<G>(s_acc, 0, aa, j, a0, a1);
                lat_mirror_update<G>(s_acc, 1, aa, j, b0, b1);
            }
            wave_sync();
        }
    }
    // the last two requests (re-reads of valid rows) must land before the LDS is released
    __builtin_amdgcn_s_waitcnt(0);
    if (active) boot_epilogue<N>(A, g, lane, s_acc);
}

// ---------------------------------------------------------------------------------------------
// Split-FP32 path.  Each key coefficient K = Khi * 2^s + Klo with Klo in [-2^(s-1), 2^(s-1));
// both halves are stored as FP32 transforms.  Digits (small integers, exact in FP32) are twisted
// in FP32, transformed and multiplied in FP32, and the hi and lo products are inverse-transformed
// separately.  Untwist and rounding run in FP64 with the 1/M scale folded into the untwist
// factors, which are re-read each step instead of held in registers; the hi product is small
// enough to round to the exact integer, so the 2^s scaling does not amplify FP32 error.
// ---------------------------------------------------------------------------------------------
// Split-FP32 key reads through a buffer resource: the per-lane part of the address is one 32-bit
// offset and the row offset is uniform (a scalar register), so no per-lane 64-bit address
// arithmetic runs in the blind-rotation loop; reads past the end of the key return zero.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t key_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float4 key_load(__amdgpu_buffer_rsrc_t rs, uint32_t lane_off, uint32_t row_off) {
    return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rs, (int)lane_off, (int)row_off, 0));
}

// kr: the key buffer; row0: byte offset of this step's row set in it.  MAC = 0: the products are
// accumulated with plain FP32 FMAs; MAC = 1: compensated (cmac_comp), with the compensation terms
// added into each sum before its inverse transform.
template <int N, int L, int BG, bool ROT, int MAC = 0>
__device__ __forceinline__ void step32(int32_t* s_acc, C2<float>* s_buf, int a, __amdgpu_buffer_rsrc_t kr,
                                       uint32_t row0, const DitTw<float>& tw, const C2<float>* tws,
                                       const FwdMfmaA& am, const C2<double>* __restrict__ untw, int lane,
                                       int split_bits, double* raw, double* raw_hi) {
    constexpr int M = N / 2, R1 = M / 64, NC = 2 * R1;
    static_assert(M == 512, "the split path runs the 512-point decimation-in-time transform");
    constexpr uint32_t OFF = decomp_offset<L, BG>();
    C2<float> h0[R1], l0[R1], h1[R1], l1[R1];
    // compensation terms of the MAC = 1 form (unused, and removed by the compiler, when MAC = 0)
    C2<float> ch0[R1], cl0[R1], ch1[R1], cl1[R1];
#pragma unroll
    for (int r = 0; r < R1; ++r) {
        h0[r] = C2<float>{0.f, 0.f}; l0[r] = C2<float>{0.f, 0.f};
        h1[r] = C2<float>{0.f, 0.f}; l1[r] = C2<float>{0.f, 0.f};
        ch0[r] = C2<float>{0.f, 0.f}; cl0[r] = C2<float>{0.f, 0.f};
        ch1[r] = C2<float>{0.f, 0.f}; cl1[r] = C2<float>{0.f, 0.f};
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        uint32_t u[NC];
#if K32_BLAYOUT
        if constexpr (ROT) {
            lat_rotate_b<lat_mirror<K32_MIRROR_G>(), OFF, !DIGIT_PERM>(u, s_acc, q, a, lane);
        } else {
#pragma unroll
            for (int r = 0; r < NC; ++r)
                u[r] = ((uint32_t)s_acc[q * N + bl_lane(lane) + bl_off(r)] + OFF) ^ (DIGIT_PERM ? 0u : OFF);
        }
#if DIGIT_PERM
        // the digit fields of each register pair, byte-permuted once for the three levels
        static_assert(BG == 7 && L == 3, "the byte layout of digits7_f16x2");
#pragma unroll
        for (int aa = 0; aa < R1; ++aa) digits7_prep(u[aa], u[aa + R1], u[aa], u[aa + R1]);
#endif
#else
        if constexpr (ROT) {
            lat_rotate<K32_MIRROR_G, OFF>(u, s_acc, q, a, lane);
        } else {
#pragma unroll
            for (int r = 0; r < NC; ++r) u[r] = ((uint32_t)s_acc[q * N + lane + 64 * r] + OFF) ^ OFF;
        }
#endif
#pragma unroll
        for (int lvl = 0; lvl < L; ++lvl) {
            C2<float> v[R1];
#if TFHE_FWD_MFMA
            // tws: the pass-2 table with the lane twist folded in (fft512_fwd_mfma)
            uint32_t dw[R1];
#pragma unroll
            for (int aa = 0; aa < R1; ++aa)
#if K32_BLAYOUT && DIGIT_PERM
                dw[aa] = digits7_f16x2(u[aa], u[aa + R1], lvl);
#else
                dw[aa] = pack_f16x2((float)digit_sx<BG>(u[aa], lvl), (float)digit_sx<BG>(u[aa + R1], lvl));
#endif
            fft512_fwd_mfma<K32_BLAYOUT>(v, dw, am, tws, tw, s_buf, lane);
#else
#pragma unroll
            for (int aa = 0; aa < R1; ++aa)
                v[aa] = C2<float>{(float)digit_sx<BG>(u[aa], lvl), (float)digit_sx<BG>(u[aa + R1], lvl)};
            fft512_dit<+1, true>(v, tw, tws, s_buf, lane);
#endif
            const uint32_t ro = row0 + (uint32_t)((q * L + lvl) * 2 * M * sizeof(float4));
#pragma unroll
            for (int r = 0; r < R1; ++r) {
                const float4 k0 = key_load(kr, (uint32_t)lane * sizeof(float4), ro + r * 64 * sizeof(float4));
                if constexpr (MAC == 0) {
                    cmac(h0[r], v[r], C2<float>{k0.x, k0.y});
                    cmac(l0[r], v[r], C2<float>{k0.z, k0.w});
                } else {
                    cmac_comp(h0[r], ch0[r], v[r], C2<float>{k0.x, k0.y});
                    cmac_comp(l0[r], cl0[r], v[r], C2<float>{k0.z, k0.w});
                }
            }
#pragma unroll
            for (int r = 0; r < R1; ++r) {
                const float4 k1 = key_load(kr, (uint32_t)lane * sizeof(float4), ro + (M + r * 64) * sizeof(float4));
                if constexpr (MAC == 0) {
                    cmac(h1[r], v[r], C2<float>{k1.x, k1.y});
                    cmac(l1[r], v[r], C2<float>{k1.z, k1.w});
                } else {
                    cmac_comp(h1[r], ch1[r], v[r], C2<float>{k1.x, k1.y});
                    cmac_comp(l1[r], cl1[r], v[r], C2<float>{k1.z, k1.w});
                }
            }
        }
    }
    if constexpr (MAC != 0) {
#pragma unroll
        for (int r = 0; r < R1; ++r) {
            h0[r] = h0[r] + ch0[r]; l0[r] = l0[r] + cl0[r];
            h1[r] = h1[r] + ch1[r]; l1[r] = l1[r] + cl1[r];
        }
    }
    fft512_dit<-1, false>(h0, tw, (const C2<float>*)nullptr, s_buf, lane);
    fft512_dit<-1, false>(l0, tw, (const C2<float>*)nullptr, s_buf, lane);
    fft512_dit<-1, false>(h1, tw, (const C2<float>*)nullptr, s_buf, lane);
    fft512_dit<-1, false>(l1, tw, (const C2<float>*)nullptr, s_buf, lane);
    const double hs = ldexp(1.0, split_bits);
    // the untwist factors are re-read every step (an opaque offset keeps the compiler from
    // hoisting them out of the blind-rotation loop into 32 long-lived registers)
    int uoff = lane * R1;
    asm volatile("" : "+v"(uoff));
#pragma unroll
    for (int aa = 0; aa < R1; ++aa) {
        const int j = lane + 64 * aa;
        const C2<double> w = untw[uoff + aa];
        C2<double> zh0 = cmulc(C2<double>{(double)h0[aa].x, (double)h0[aa].y}, w);
        C2<double> zl0 = cmulc(C2<double>{(double)l0[aa].x, (double)l0[aa].y}, w);
        C2<double> zh1 = cmulc(C2<double>{(double)h1[aa].x, (double)h1[aa].y}, w);
        C2<double> zl1 = cmulc(C2<double>{(double)l1[aa].x, (double)l1[aa].y}, w);
        const uint32_t r00 = (round_mod32(zh0.x) << split_bits) + round_mod32(zl0.x);
        const uint32_t r01 = (round_mod32(zh0.y) << split_bits) + round_mod32(zl0.y);
        const uint32_t r10 = (round_mod32(zh1.x) << split_bits) + round_mod32(zl1.x);
        const uint32_t r11 = (round_mod32(zh1.y) << split_bits) + round_mod32(zl1.y);
        if (raw) {
            raw[j] = zh0.x * hs + zl0.x; raw[j + M] = zh0.y * hs + zl0.y;
            raw[N + j] = zh1.x * hs + zl1.x; raw[N + j + M] = zh1.y * hs + zl1.y;
        }
        if (raw_hi) {   // the hi product before rounding, in units of 2^split_bits
            raw_hi[j] = zh0.x; raw_hi[j + M] = zh0.y;
            raw_hi[N + j] = zh1.x; raw_hi[N + j + M] = zh1.y;
        }
        if (ROT) {
            const int32_t a0 = (int32_t)((uint32_t)s_acc[j] + r00);
            const int32_t a1 = (int32_t)((uint32_t)s_acc[j + M] + r01);
            const int32_t b0 = (int32_t)((uint32_t)s_acc[N + j] + r10);
            const int32_t b1 = (int32_t)((uint32_t)s_acc[N + j + M] + r11);
            s_acc[j] = a0;
            s_acc[j + M] = a1;
            s_acc[N + j] = b0;
            s_acc[N + j + M] = b1;
            lat_mirror_update<K32_MIRROR_G>(s_acc, 0, aa, j, a0, a1);
            lat_mirror_update<K32_MIRROR_G>(s_acc, 1, aa, j, b0, b1);
        } else {
            s_acc[j] = (int32_t)r00; s_acc[j + M] = (int32_t)r01;
            s_acc[N + j] = (int32_t)r10; s_acc[N + j + M] = (int32_t)r11;
        }
    }
    wave_sync();
}

// Fused keyswitch (KSF): the wavefront keyswitches its own extracted sample straight from the
// accumulator in LDS instead of writing the sample out for k_keyswitch.  Extraction: u_0 = a_0,
// u_j = -a_{N-j}, u_N = b_0; keyswitch as k_keyswitch (digit pairs, lwe-keyswitch-functions.cpp:
// 114-129, 228-241), bit-identical.  Lane L owns output coefficients CPL*L .. CPL*L+CPL-1 of the
// NP-padded rows, read with 8-byte loads; the digit pair is uniform, and a zero pair reads row 0
// under a zero mask, so the loop has no branches and its loads can be issued ahead.
template <int TKS, int KN, int NP>
__device__ __forceinline__ void keyswitch_from_acc(const int32_t* s_acc, const int32_t* __restrict__ ksk,
                                                   int32_t* __restrict__ out, int n, int lane, int32_t add_b) {
    constexpr int TP = TKS / 2, RPI = TP * KS_PAIR_ROWS, CPL = NP / 64;
    static_assert(NP % 128 == 0 || CPL % 2 == 0, "8-byte row reads");
    const uint32_t prec_off = 1u << (32 - (1 + 2 * TKS));
    uint32_t acc[CPL];
#pragma unroll
    for (int t = 0; t < CPL; ++t) acc[t] = 0;
    const int32_t* kl = ksk + (size_t)CPL * lane;
#pragma unroll 2
    for (int i = 0; i < KN; ++i) {
        const uint32_t ui = (i == 0) ? (uint32_t)s_acc[0] : 0u - (uint32_t)s_acc[KN - i];
        const uint32_t abar = (uint32_t)__builtin_amdgcn_readfirstlane((int)(ui + prec_off));
#pragma unroll
        for (int jp = 0; jp < TP; ++jp) {
            const uint32_t d = (abar >> (28 - 4 * jp)) & 15u;
            const uint32_t m = d ? 0xFFFFFFFFu : 0u;
            const int row = d ? jp * KS_PAIR_ROWS + (int)d - 1 : 0;
            const int2* kr = (const int2*)(kl + ((size_t)i * RPI + row) * NP);
#pragma unroll
            for (int t = 0; t < CPL / 2; ++t) {
                const int2 v = kr[t];
                acc[2 * t] -= (uint32_t)v.x & m;
                acc[2 * t + 1] -= (uint32_t)v.y & m;
            }
        }
    }
    const uint32_t bterm = (uint32_t)s_acc[KN] + (uint32_t)add_b;
#pragma unroll
    for (int t = 0; t < CPL; ++t) {
        const int c = CPL * lane + t;
        if (c <= n) out[c] = (int32_t)(acc[t] + (c == n ? bterm : 0u));
    }
}

// MAC = 1 (compensated accumulation) doubles the accumulator registers and is compiled for one
// wavefront per SIMD.  KSF: fused keyswitch (mode 0, one segment, N = 1024 set), output to
// L.ks_out[g] instead of an extracted sample at L.u_out[g].
template <int N, int L, int BG, int MAC = 0, bool KSF = false>
__global__ void __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(N <= 1024 && MAC == 0 ? TFHE_BOOT32_WAVES : 1,
                                   N <= 1024 && MAC == 0 ? TFHE_BOOT32_WAVES : 1))) k_boot32(BootKArgs A) {
    constexpr int M = N / 2, R1 = M / 64, KPL = 2 * L;
    // the accumulator with mirrored ends (lat_rotate); the modulus-switched rotations are read 64
    // steps at a time from the inputs (one per lane, v_readlane per step) instead of from LDS
    constexpr int MW = lat_mirror<K32_MIRROR_G>();
    __shared__ int32_t s_acc_m[2 * N + 2 * MW];
    __shared__ C2<float> s_buf[xbuf_elems<M>()];
    int32_t* const s_acc = s_acc_m + MW;
    const int lane = threadIdx.x;
    const int g = blockIdx.x;
    boot_prologue<N>(A, g, lane, s_acc, nullptr);
    lat_mirror_fill<K32_MIRROR_G>(s_acc, lane);
    wave_sync();

    DitTw<float> tw;
    load_dit_tw(tw, (const C2<float>*)A.t2, (const C2<float>*)A.t1, lane);
    // forward twist (fft512_dit), or the matrix-core form's pass-2 table and matrix pieces
    C2<float> tws[R1];
    FwdMfmaA am;
#if TFHE_FWD_MFMA
#pragma unroll
    for (int r = 0; r < R1; ++r) tws[r] = ((const C2<float>*)A.fwd_t2)[lane * R1 + r];
    load_fwd_mfma_a(am, (const f16x8_t*)A.fwd_a16, lane);
#else
#pragma unroll
    for (int r = 0; r < R1; ++r) tws[r] = ((const C2<float>*)A.twist32)[lane * R1 + r];
#endif
    const C2<double>* untw = (const C2<double>*)A.untw64;
    // one row set (2L rows of 2 polynomials of M float4) per LWE coefficient; the whole key is far
    // below 4 GiB, so byte offsets fit 32 bits
    constexpr uint32_t ROWSET_B = (uint32_t)(KPL * 2 * M * sizeof(float4));
    const __amdgpu_buffer_rsrc_t key = key_rsrc(A.bk, (uint32_t)A.n * ROWSET_B);
    const int sb = A.split_bits;

    if (A.L.mode == 2) {
        double* raw = A.L.raw_out ? A.L.raw_out + (size_t)g * 2 * N : nullptr;
        double* raw_hi = A.L.raw_hi ? A.L.raw_hi + (size_t)g * 2 * N : nullptr;
        step32<N, L, BG, false, MAC>(s_acc, s_buf, 0, key, (uint32_t)A.L.row_set[g] * ROWSET_B, tw, tws, am, untw, lane,
                                     sb, raw, raw_hi);
    } else {
        const int n = A.n;
        const BootIn in = A.L.mode != 1 ? boot_inputs(A, g) : BootIn{nullptr, nullptr, 0u, 0u, 0};
        auto bara_at = [&](int i) -> int {   // rotation of step i, 0 past the last step
            if (i >= n) return 0;
            if (A.L.mode == 1) return A.L.bara_in[(size_t)g * n + i];
            uint32_t x = in.al * (uint32_t)in.xa[i];
            if (in.xb) x += in.be * (uint32_t)in.xb[i];
            return mod_switch_2n<N>(x);
        };
        int bcur = bara_at(lane), bnext = bara_at(64 + lane);
        for (int i = 0; i < n; ++i) {
            if ((i & 63) == 0 && i > 0) {
                bcur = bnext;
                bnext = bara_at(i + 64 + lane);
            }
            const int a = __builtin_amdgcn_readlane(bcur, i & 63);
            if (a == 0) continue;
            step32<N, L, BG, true, MAC>(s_acc, s_buf, a, key, (uint32_t)i * ROWSET_B, tw, tws, am, untw, lane, sb,
                                        nullptr, nullptr);
        }
    }
    if constexpr (KSF) {
        keyswitch_from_acc<8, N, 640>(s_acc, A.ksk, A.L.ks_out + (size_t)g * (A.n + 1), A.n, lane, A.L.ks_add_b);
    } else {
        boot_epilogue<N>(A, g, lane, s_acc);
    }
}

// ---------------------------------------------------------------------------------------------
// Split-FP32 path, latency form: one bootstrap per workgroup of 2l wavefronts, as k_boot64lat.
// Wavefront w = (q, lvl) transforms its digits and leaves four partial products in LDS (hi and lo
// key halves of both output polynomials); after a barrier, wavefronts 0-3 each sum the 2l partials
// of one (output polynomial, half), invert it, untwist and round it in FP64, and add it into the
// accumulator with an LDS integer atomic (the hi half shifted by the split), exact mod 2^32 in
// either order.
// ---------------------------------------------------------------------------------------------
template <int N, int L, int BG>
__global__ void __launch_bounds__(64 * 2 * L) k_boot32lat(BootKArgs A) {
    constexpr int M = N / 2, R1 = M / 64, NC = 2 * R1, KPL = 2 * L;
    static_assert(M == 512 && KPL >= 4, "the 512-point decimation-in-time transform; four output halves");
    constexpr uint32_t OFF = decomp_offset<L, BG>();
    constexpr uint32_t ROW_B = (uint32_t)(2 * M * sizeof(float4)), ROWSET_B = KPL * ROW_B;
    // accumulator with mirrored ends: 128 words for the operand-layout rotation (lat_rotate_b)
    // that the matrix-core forward reads, 448 for lat_rotate
    constexpr int MG = LAT32_BLAYOUT ? 3 : 8, MW = lat_mirror<MG>();
    __shared__ int32_t s_acc_m[2 * N + 2 * MW];
    int32_t* const s_acc = s_acc_m + MW;
    __shared__ int16_t s_bara[kMaxLweN];
    __shared__ C2<float> s_x[KPL][xbuf_elems<M>()];
    __shared__ C2<float> s_part[KPL][4][M];   // [key row][(output polynomial j, half h) = 2j + h]
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int g = blockIdx.x;
    if (w == 0) {
        boot_prologue<N>(A, g, lane, s_acc, s_bara);
        lat_mirror_fill<MG>(s_acc, lane);
    }
    __syncthreads();

    DitTw<float> tw;
    load_dit_tw(tw, (const C2<float>*)A.t2, (const C2<float>*)A.t1, lane);
    // forward twist (fft512_dit), or the matrix-core form's pass-2 table and matrix pieces
    C2<float> tws[R1];
    FwdMfmaA am;
#if TFHE_LAT_FWD_MFMA
#pragma unroll
    for (int r = 0; r < R1; ++r) tws[r] = ((const C2<float>*)A.fwd_t2)[lane * R1 + r];
    load_fwd_mfma_a(am, (const f16x8_t*)A.fwd_a16, lane);
#else
#pragma unroll
    for (int r = 0; r < R1; ++r) tws[r] = ((const C2<float>*)A.twist32)[lane * R1 + r];
#endif
    const C2<double>* untw = (const C2<double>*)A.untw64;
    const __amdgpu_buffer_rsrc_t key = key_rsrc(A.bk, (uint32_t)A.n * ROWSET_B);
    const int sb = A.split_bits;
    const int q = w / L, lvl = w % L;
    float4 k0[R1], k1[R1];
    auto load_row = [&](int i) {
        const uint32_t ro = (uint32_t)i * ROWSET_B + (uint32_t)w * ROW_B;
#pragma unroll
        for (int r = 0; r < R1; ++r) {
            k0[r] = key_load(key, (uint32_t)lane * sizeof(float4), ro + r * 64 * sizeof(float4));
            k1[r] = key_load(key, (uint32_t)lane * sizeof(float4), ro + (M + r * 64) * sizeof(float4));
        }
    };
    const int n = A.n;
    int loaded = -1;
    for (int i = 0; i < n; ++i) {
        const int a = __builtin_amdgcn_readfirstlane((int)s_bara[i]);
        if (a == 0) continue;
        if (loaded != i) load_row(i);
        {
            uint32_t u[NC];
            // with the matrix-core forward and DIGIT_PERM the rotation leaves x + offset for
            // digits7_f16x2 (the pairs (u[aa], u[aa + R1]) are the same in either layout)
            constexpr bool XO = !(TFHE_LAT_FWD_MFMA && DIGIT_PERM);
#if LAT32_BLAYOUT
            lat_rotate_b<MW, OFF, XO>(u, s_acc, q, a, lane);
#else
            lat_rotate<MG, OFF, XO>(u, s_acc, q, a, lane);
#endif
            C2<float> v[R1];
#if TFHE_LAT_FWD_MFMA
            uint32_t dw[R1];
#if DIGIT_PERM
            static_assert(BG == 7 && L == 3, "the byte layout of digits7_f16x2");
#pragma unroll
            for (int aa = 0; aa < R1; ++aa) dw[aa] = digits7_f16x2_one(u[aa], u[aa + R1], lvl);
#else
#pragma unroll
            for (int aa = 0; aa < R1; ++aa)
                dw[aa] = pack_f16x2((float)digit_sx<BG>(u[aa], lvl), (float)digit_sx<BG>(u[aa + R1], lvl));
#endif
            fft512_fwd_mfma<LAT32_BLAYOUT>(v, dw, am, tws, tw, s_x[w], lane);
#else
#pragma unroll
            for (int aa = 0; aa < R1; ++aa)
                v[aa] = C2<float>{(float)digit_sx<BG>(u[aa], lvl), (float)digit_sx<BG>(u[aa + R1], lvl)};
            fft512_dit<+1, true>(v, tw, tws, s_x[w], lane);
#endif
#pragma unroll
            for (int r = 0; r < R1; ++r) {
                s_part[w][0][r * 64 + lane] = cmul(v[r], C2<float>{k0[r].x, k0[r].y});
                s_part[w][1][r * 64 + lane] = cmul(v[r], C2<float>{k0[r].z, k0[r].w});
                s_part[w][2][r * 64 + lane] = cmul(v[r], C2<float>{k1[r].x, k1[r].y});
                s_part[w][3][r * 64 + lane] = cmul(v[r], C2<float>{k1[r].z, k1[r].w});
            }
        }
        if (i + 1 < n) load_row(i + 1);
        loaded = i + 1;
        __syncthreads();
        if (w < 4) {
            const int j = w >> 1, hi = !(w & 1);
            C2<float> o[R1];
#pragma unroll
            for (int r = 0; r < R1; ++r) {
                o[r] = s_part[0][w][r * 64 + lane];
#pragma unroll
                for (int t = 1; t < KPL; ++t) o[r] = o[r] + s_part[t][w][r * 64 + lane];
            }
            fft512_dit<-1, false>(o, tw, (const C2<float>*)nullptr, s_x[w], lane);
            int uoff = lane * R1;
            asm volatile("" : "+v"(uoff));
            int32_t* acc = s_acc + j * N;
            const int sh = hi ? sb : 0;
#pragma unroll
            for (int aa = 0; aa < R1; ++aa) {
                const int jj = lane + 64 * aa;
                const C2<double> z = cmulc(C2<double>{(double)o[aa].x, (double)o[aa].y}, untw[uoff + aa]);
                const int32_t dl = (int32_t)(round_mod32(z.x) << sh), dh = (int32_t)(round_mod32(z.y) << sh);
                atomicAdd(acc + jj, dl);
                atomicAdd(acc + jj + M, dh);
                // the mirrors take the same increments: p0 from N - MW up, p1 below MW
                if (j == 0 && aa >= 8 - MW / 64) atomicAdd(acc + jj + M - N, dh);
                if (j == 1 && aa < MW / 64) atomicAdd(acc + jj + N, dl);
            }
        }
        __syncthreads();
    }
    if (w == 0) boot_epilogue<N>(A, g, lane, s_acc);
}

// ---------------------------------------------------------------------------------------------
// Keyswitch (lwe-keyswitch-functions.cpp:114-129, :228-241), bit-exact.
// The t base-4 digits of each input coefficient are taken in pairs: the key holds, for every
// input index i and digit pair g, the 15 sums ks[i][2g][d1] + ks[i][2g+1][d2] over (d1, d2) != 0
// (ks[.][.][0] = 0, the row the reference skips), so one row subtraction replaces two.  Integer
// sums wrap mod 2^32 as the reference's two subtractions do, so the result is bit-identical.
// Workgroup = KS_GATES ciphertexts (one per lane) x a chunk of C output coefficients (registers).
// For each input index i the workgroup stages the (t/2)*15 rows of the chunk in LDS, with a zero
// row for the pair (0, 0), and every lane subtracts the row its own 4-bit digit pair selects: one
// LDS read and one subtract per (ciphertext, i, pair, coefficient), each key row read from memory
// once per workgroup.  The extracted samples are staged through LDS in blocks of KS_IBLK indices
// so their reads stay coalesced.
// ---------------------------------------------------------------------------------------------
constexpr int KS_GATES = 256;
#ifndef TFHE_KS_C
#define TFHE_KS_C 32   // output coefficients per workgroup (registers per lane)
#endif
// Occupancy: 8 input indices per staging block (26 KB of LDS per workgroup at t = 8) and registers
// for 6 wavefronts per SIMD let six workgroups share a CU, so a 16384-ciphertext launch (1280
// workgroups) is resident at once instead of running 1.25 rounds of four per CU.
#ifndef TFHE_KS_IBLK
#define TFHE_KS_IBLK 8    // input indices per staging block of the extracted samples
#endif
constexpr int KS_IBLK = TFHE_KS_IBLK;
#ifndef TFHE_KS_WPE
#define TFHE_KS_WPE 6     // wavefronts per SIMD the register allocation must allow (t = 8; at t = 10
#endif                    // the kernel needs 104 VGPRs, so it asks for 4)

template <int TKS, int KN, int NP, int C>
__global__ void __launch_bounds__(KS_GATES) __attribute__((amdgpu_waves_per_eu(TKS == 8 ? TFHE_KS_WPE : 4))) k_keyswitch(const int32_t* __restrict__ u1, const int32_t* __restrict__ u2,
                                                        int32_t add_b, const int32_t* __restrict__ ksk,
                                                        int32_t* __restrict__ out, int count, int n,
                                                        int32_t* const* __restrict__ out_tab, int out_seg) {
    static_assert(KS_GATES % C == 0 && NP % C == 0 && KN % KS_IBLK == 0 && TKS % 2 == 0, "keyswitch tiling");
    constexpr int TP = TKS / 2;                      // digit pairs
    constexpr int RPI = TP * KS_PAIR_ROWS;           // key rows per input index
    // rows padded by one uint4 (36 dwords apart): the 16 rows of a pair start on distinct bank quads
    constexpr int KROW = C / 4 + 1;
    __shared__ uint4 s_k[2][TP][16][KROW];
    __shared__ uint32_t s_u[KS_IBLK][KS_GATES];
    constexpr size_t UST = (size_t)KN + 1;
    constexpr int KV = RPI * C;                      // staged values per input index
    constexpr int KPT = (KV + KS_GATES - 1) / KS_GATES;
    constexpr int ROWS_PER_T = KS_GATES / C;         // rows covered by one pass of the block
    const int tid = threadIdx.x;
    const int g0 = blockIdx.y * KS_GATES;
    const int g = g0 + tid;
    const int c0 = blockIdx.x * C;
    const uint32_t prec_off = 1u << (32 - (1 + 2 * TKS));

    uint32_t acc[C];
#pragma unroll
    for (int c = 0; c < C; ++c) acc[c] = 0;
    for (int idx = tid; idx < 2 * TP * (C / 4); idx += KS_GATES) {
        const int b = idx / (TP * (C / 4)), r = idx % (TP * (C / 4));
        s_k[b][r / (C / 4)][0][r % (C / 4)] = make_uint4(0, 0, 0, 0);
    }
    // staged value v = tid + t * KS_GATES covers row jh = v / C = tid / C + t * ROWS_PER_T and
    // column cc = tid % C of the chunk
    const int cc = tid % C, jh0 = tid / C;
    const int32_t* kbase = ksk + (size_t)jh0 * NP + c0 + cc;
    uint32_t* sk_flat = reinterpret_cast<uint32_t*>(&s_k[0][0][0][0]);
    uint32_t pre[KPT];
    auto load_rows = [&](int i) {
#pragma unroll
        for (int t = 0; t < KPT; ++t) {
            const int jh = jh0 + t * ROWS_PER_T;
            pre[t] = (jh < RPI) ? (uint32_t)kbase[((size_t)i * RPI + t * ROWS_PER_T) * NP] : 0u;
        }
    };
    auto store_rows = [&](int b) {
#pragma unroll
        for (int t = 0; t < KPT; ++t) {
            const int jh = jh0 + t * ROWS_PER_T;
            if (jh < RPI)
                sk_flat[((b * TP + jh / KS_PAIR_ROWS) * 16 + jh % KS_PAIR_ROWS + 1) * (KROW * 4) + cc] = pre[t];
        }
    };
    load_rows(0);
    store_rows(0);
    // u staging: thread reads row gg = tid / KS_IBLK + t * (KS_GATES / KS_IBLK), column ii = tid % KS_IBLK
    constexpr int GSTEP = KS_GATES / KS_IBLK;
    const int ii0 = tid % KS_IBLK, gg0 = tid / KS_IBLK;
    const int32_t* u1b = u1 + (size_t)(g0 + gg0) * UST + ii0;
    const int32_t* u2b = u2 ? u2 + (size_t)(g0 + gg0) * UST + ii0 : nullptr;
    for (int i0 = 0; i0 < KN; i0 += KS_IBLK) {
        __syncthreads();
#pragma unroll
        for (int t = 0; t < KS_IBLK; ++t) {
            const int gg = gg0 + t * GSTEP;
            uint32_t x = 0;
            if (g0 + gg < count) {
                x = (uint32_t)u1b[(size_t)t * GSTEP * UST + i0];
                if (u2b) x += (uint32_t)u2b[(size_t)t * GSTEP * UST + i0];
            }
            s_u[ii0][gg] = x + prec_off;
        }
        __syncthreads();
#pragma unroll 1
        for (int ii = 0; ii < KS_IBLK; ++ii) {
            const int i = i0 + ii;
            const int b = i & 1;
            if (i + 1 < KN) load_rows(i + 1);
            const uint32_t aibar = s_u[ii][tid];
            // not unrolled: an unrolled pair loop lets the compiler re-associate the row
            // subtractions into one sum and keep every row live
#pragma unroll 1
            for (int jp = 0; jp < TP; ++jp) {
                // digits 2jp (high two bits) and 2jp+1: (aibar >> (32 - 2(2jp+2))) & 15
                const uint32_t d = (aibar >> (28 - 4 * jp)) & 15u;
                const uint4* row = &s_k[b][jp][d][0];
#pragma unroll
                for (int c4 = 0; c4 < C / 4; ++c4) {
                    const uint4 kv = row[c4];
                    acc[4 * c4 + 0] -= kv.x;
                    acc[4 * c4 + 1] -= kv.y;
                    acc[4 * c4 + 2] -= kv.z;
                    acc[4 * c4 + 3] -= kv.w;
                }
            }
            if (i + 1 < KN) store_rows(b ^ 1);
            __syncthreads();
        }
    }
    if (g < count) {
        if (n >= c0 && n < c0 + C) {
            uint32_t bterm = (uint32_t)u1[(size_t)g * UST + KN] + (uint32_t)add_b;
            if (u2) bterm += (uint32_t)u2[(size_t)g * UST + KN];
#pragma unroll
            for (int c = 0; c < C; ++c)
                if (c0 + c == n) acc[c] += bterm;
        }
        // output g goes to out[g], or to sample g mod out_seg of the array out_tab[g / out_seg]
        int32_t* o = out_tab ? out_tab[g / out_seg] + (size_t)(g % out_seg) * (n + 1) : out + (size_t)g * (n + 1);
#pragma unroll
        for (int c = 0; c < C; ++c)
            if (c0 + c <= n) o[c0 + c] = (int32_t)acc[c];
    }
}

// ---------------------------------------------------------------------------------------------
// Keyswitch of a few ciphertexts (latency).  The tiled kernel above walks all KN input indices in
// sequence in each of its NP/C workgroups, which for a single gate is most of a millisecond.  Here
// the input indices are spread over KS_SLICES workgroups per ciphertext and 64-coefficient chunk:
// each subtracts its slice's selected digit-pair rows straight from the key (the digit pair is
// uniform in the workgroup, the read coalesced across coefficients), and adds its partial sum into
// the output with a 32-bit atomic.  Integer addition wraps mod 2^32 in any order, so the result
// is bit-identical to the tiled kernel's.  k_ks_clear zeroes the outputs first (a kernel rather
// than a memset, so that a captured graph records no memset node).
// ---------------------------------------------------------------------------------------------
constexpr int KS_SLICES = 32;
// ciphertexts per launch up to which the sliced form is used: its time grows with the count (35 us
// for 1, 120 us for 64), the tiled kernel's does not below 2048 (1.2 ms)
constexpr int KS_SMALL_MAX = 512;

__device__ __forceinline__ int32_t* ks_out_row(int32_t* out, int32_t* const* out_tab, int out_seg, int g, int n) {
    return out_tab ? out_tab[g / out_seg] + (size_t)(g % out_seg) * (n + 1) : out + (size_t)g * (n + 1);
}

__global__ void __launch_bounds__(256) k_ks_clear(int32_t* out, int n, int32_t* const* out_tab, int out_seg) {
    int32_t* o = ks_out_row(out, out_tab, out_seg, blockIdx.x, n);
    for (int c = threadIdx.x; c <= n; c += 256) o[c] = 0;
}

template <int TKS, int KN, int NP>
__global__ void __launch_bounds__(64) k_keyswitch_sliced(const int32_t* __restrict__ u1,
                                                         const int32_t* __restrict__ u2, int32_t add_b,
                                                         const int32_t* __restrict__ ksk, int32_t* out, int n,
                                                         int32_t* const* out_tab, int out_seg) {
    static_assert(KN % KS_SLICES == 0 && NP % 64 == 0 && TKS % 2 == 0, "sliced keyswitch tiling");
    constexpr int TP = TKS / 2, RPI = TP * KS_PAIR_ROWS, ISL = KN / KS_SLICES;
    constexpr size_t UST = (size_t)KN + 1;
    const int c = blockIdx.x * 64 + threadIdx.x;   // < NP: the key rows are padded to NP
    const int s = blockIdx.y;
    const int g = blockIdx.z;
    const uint32_t prec_off = 1u << (32 - (1 + 2 * TKS));
    const int32_t* x1 = u1 + (size_t)g * UST;
    const int32_t* x2 = u2 ? u2 + (size_t)g * UST : nullptr;
    uint32_t acc = 0;
    for (int i = s * ISL; i < (s + 1) * ISL; ++i) {
        uint32_t abar = (uint32_t)x1[i] + prec_off;
        if (x2) abar += (uint32_t)x2[i];
        const int32_t* kr = ksk + (size_t)i * RPI * NP + c;
#pragma unroll
        for (int jp = 0; jp < TP; ++jp) {
            const uint32_t d = (abar >> (28 - 4 * jp)) & 15u;
            if (d) acc -= (uint32_t)kr[(size_t)(jp * KS_PAIR_ROWS + (int)d - 1) * NP];
        }
    }
    if (c > n) return;
    if (s == 0 && c == n) {
        acc += (uint32_t)x1[KN] + (uint32_t)add_b;
        if (x2) acc += (uint32_t)x2[KN];
    }
    atomicAdd(ks_out_row(out, out_tab, out_seg, g, n) + c, (int32_t)acc);
}

// ---------------------------------------------------------------------------------------------
// Key preparation: one wavefront per key polynomial.  Both outputs come from the FP64 transform
// (twist folded into the pass-1 twiddles), so the FP32 split key is the correctly rounded FP64
// transform of each exact integer half.
// ---------------------------------------------------------------------------------------------
template <int N>
__global__ void __launch_bounds__(64) k_keyprep(const int32_t* __restrict__ bk, C2<double>* __restrict__ out64,
                                                float4* __restrict__ out32, int split_bits,
                                                const C2<double>* __restrict__ t1, const C2<double>* __restrict__ t2,
                                                const C2<double>* __restrict__ wa64) {
    constexpr int M = N / 2, R1 = M / 64;
    __shared__ C2<double> s_buf[xbuf_elems<M>()];
    const int lane = threadIdx.x;
    const size_t poly = blockIdx.x;
    const int32_t* p = bk + poly * N;
    LaneTw<double, R1> tw;
    load_lane_tw(tw, t1, t2, lane);
    C2<double> wa[R1];
#pragma unroll
    for (int r = 0; r < R1; ++r) wa[r] = wa64[r];
    C2<double> v[R1];
    if (out64) {
#pragma unroll
        for (int aa = 0; aa < R1; ++aa) {
            v[aa] = C2<double>{(double)p[lane + 64 * aa], (double)p[lane + 64 * aa + M]};
            v[aa] = cmul(v[aa], wa[aa]);
        }
        fft_fwd<R1>(v, tw, s_buf, lane);
#pragma unroll
        for (int r = 0; r < R1; ++r) out64[poly * M + r * 64 + lane] = v[r];
    }
    if (out32) {
        C2<double> h[R1];
        const int64_t half = (int64_t)1 << (split_bits - 1);
#pragma unroll
        for (int aa = 0; aa < R1; ++aa) {
            const int64_t x0 = p[lane + 64 * aa], x1 = p[lane + 64 * aa + M];
            const int64_t hi0 = (x0 + half) >> split_bits, hi1 = (x1 + half) >> split_bits;
            const int64_t lo0 = x0 - (hi0 << split_bits), lo1 = x1 - (hi1 << split_bits);
            h[aa] = cmul(C2<double>{(double)hi0, (double)hi1}, wa[aa]);
            v[aa] = cmul(C2<double>{(double)lo0, (double)lo1}, wa[aa]);
        }
        fft_fwd<R1>(h, tw, s_buf, lane);
        fft_fwd<R1>(v, tw, s_buf, lane);
#pragma unroll
        for (int r = 0; r < R1; ++r)
            out32[poly * M + r * 64 + lane] = make_float4((float)h[r].x, (float)h[r].y, (float)v[r].x, (float)v[r].y);
    }
}

__global__ void k_ksk_relayout(const int32_t* __restrict__ in, int32_t* __restrict__ out, int kN, int t, int n,
                               int np) {
    // out[i][g][4 d1 + d2 - 1][c] = in[i][2g][d1][c] + in[i][2g+1][d2][c] (mod 2^32) for
    // (d1, d2) != (0, 0), with in[.][.][0] taken as zero; c <= n, zero padding for c > n
    const size_t total = (size_t)kN * (t / 2) * KS_PAIR_ROWS * np;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(idx % np);
        const size_t r = idx / np;  // (i*(t/2) + g)*15 + hh - 1
        const int hh = (int)(r % KS_PAIR_ROWS) + 1, d1 = hh >> 2, d2 = hh & 3;
        const size_t ig = r / KS_PAIR_ROWS;         // i*(t/2) + g, so digit 2g of index i is row 2*ig
        uint32_t v = 0;
        if (c <= n) {
            if (d1) v += (uint32_t)in[((2 * ig) * 4 + d1) * (size_t)(n + 1) + c];
            if (d2) v += (uint32_t)in[((2 * ig + 1) * 4 + d2) * (size_t)(n + 1) + c];
        }
        out[idx] = (int32_t)v;
    }
}

size_t ksk_device_bytes(const DevParams& p, int np) {
    return (size_t)p.N * (p.ks_t / 2) * KS_PAIR_ROWS * np * sizeof(int32_t);
}

// LWE linear jobs of a circuit level (NOT, COPY, constants): one workgroup per (desc, sample).
__global__ void __launch_bounds__(64) k_lwe_linear(const LinDesc* __restrict__ descs, int count, int n) {
    const int d = blockIdx.x / count, i = blockIdx.x - d * count;
    const LinDesc D = descs[d];
    int32_t* dst = D.dst + (size_t)i * (n + 1);
    const int32_t* src = D.src ? D.src + (size_t)i * (n + 1) : nullptr;
    for (int j = threadIdx.x; j <= n; j += 64) {
        uint32_t v = src ? (uint32_t)D.mul * (uint32_t)src[j] : 0u;
        if (j == n) v += (uint32_t)D.add_b;
        dst[j] = (int32_t)v;
    }
}

hipError_t launch_lwe_linear(const LinDesc* d_descs, int ndesc, int count, int n, hipStream_t st) {
    if (ndesc <= 0 || count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_lwe_linear, dim3(ndesc * count), dim3(64), 0, st, d_descs, count, n);
    return hipGetLastError();
}

template <int N>
__global__ void k_extract(int32_t* __restrict__ u, const int32_t* __restrict__ acc, int count) {
    const int g = blockIdx.x;
    const int32_t* a = acc + (size_t)g * 2 * N;
    int32_t* o = u + (size_t)g * (N + 1);
    for (int j = threadIdx.x; j < N; j += blockDim.x) o[j] = (j == 0) ? a[0] : (int32_t)(0u - (uint32_t)a[N - j]);
    if (threadIdx.x == 0) o[N] = a[N];
}

template <int N, int L, int BG>
__global__ void __launch_bounds__(64) k_rotdec(int32_t* __restrict__ rot, int32_t* __restrict__ digits,
                                               const int32_t* __restrict__ acc, const int32_t* __restrict__ bara) {
    constexpr uint32_t OFF = decomp_offset<L, BG>();
    __shared__ int32_t s_acc[2 * N];
    const int g = blockIdx.x, lane = threadIdx.x;
    for (int j = lane; j < 2 * N; j += 64) s_acc[j] = acc[(size_t)g * 2 * N + j];
    __syncthreads();
    // the blind-rotation loop's own forms: rotdiff_e4 and the signed bit-field digits, with the
    // plain forms (rotdiff, digit) as a cross-check that poisons the output on any disagreement
    const int a = bara[g];
    for (int q = 0; q < 2; ++q)
        for (int j = lane; j < N; j += 64) {
            const uint32_t e4 = rot_base4(lane, a) + 256u * (uint32_t)(j >> 6);
            const uint32_t t = rotdiff_e4<N>(s_acc + q * N, e4, j, 0);
            const uint32_t ux = rotdiff_e4<N>(s_acc + q * N, e4, j, OFF) ^ OFF;
            const bool same = t == rotdiff<N>(s_acc + q * N, j, a);
            rot[((size_t)g * 2 + q) * N + j] = same ? (int32_t)t : (int32_t)(t ^ 0x5a5a5a5au);
#pragma unroll
            for (int lvl = 0; lvl < L; ++lvl) {
                const int d = digit_sx<BG>(ux, lvl);
                digits[(((size_t)g * 2 + q) * L + lvl) * N + j] = d == digit<BG>(t + OFF, lvl) ? d : 1 << 20;
            }
        }
}

// natural frequency index of lane-order slot (lane, s, k3)
template <int R1>
__device__ __forceinline__ int lane_order_k(int lane, int s, int k3) {
    return (lane & 7) + 8 * s + R1 * (lane >> 3) + 8 * R1 * k3;
}

template <int N>
__global__ void __launch_bounds__(64) k_fft_test64(double* __restrict__ out, const void* __restrict__ in, int inverse,
                                                   const C2<double>* __restrict__ t1, const C2<double>* __restrict__ t2,
                                                   const C2<double>* __restrict__ wa64) {
    constexpr int M = N / 2, R1 = M / 64;
    __shared__ C2<double> s_buf[xbuf_elems<M>()];
    const int lane = threadIdx.x;
    const size_t g = blockIdx.x;
    LaneTw<double, R1> tw;
    load_lane_tw(tw, t1, t2, lane);
    C2<double> v[R1];
    if (!inverse) {
        const int32_t* p = (const int32_t*)in + g * N;
#pragma unroll
        for (int aa = 0; aa < R1; ++aa)
            v[aa] = cmul(C2<double>{(double)p[lane + 64 * aa], (double)p[lane + 64 * aa + M]}, wa64[aa]);
        fft_fwd<R1>(v, tw, s_buf, lane);
        double* o = out + g * N;
#pragma unroll
        for (int s = 0; s < R1 / 8; ++s)
#pragma unroll
            for (int k3 = 0; k3 < 8; ++k3) {
                const int k = lane_order_k<R1>(lane, s, k3);
                o[2 * k] = v[s * 8 + k3].x;
                o[2 * k + 1] = v[s * 8 + k3].y;
            }
    } else {
        const double* F = (const double*)in + g * N;
#pragma unroll
        for (int s = 0; s < R1 / 8; ++s)
#pragma unroll
            for (int k3 = 0; k3 < 8; ++k3) {
                const int k = lane_order_k<R1>(lane, s, k3);
                v[s * 8 + k3] = C2<double>{F[2 * k], F[2 * k + 1]};
            }
        fft_inv<R1>(v, tw, s_buf, lane);
        double* o = out + g * N;
#pragma unroll
        for (int aa = 0; aa < R1; ++aa) {
            C2<double> z = cmulc(v[aa], wa64[aa]);
            o[lane + 64 * aa] = z.x / M;
            o[lane + 64 * aa + M] = z.y / M;
        }
    }
}

// mode 0: forward (fft512_dit with the twist); 1: inverse; 2: forward of digit inputs (|p_j| < 2048)
// with the first pass on the matrix cores (fft512_fwd_mfma), as the split bootstrap kernel runs it
template <int N>
__global__ void __launch_bounds__(64) k_fft_test32(double* __restrict__ out, const void* __restrict__ in, int inverse,
                                                   const C2<float>* __restrict__ t1, const C2<float>* __restrict__ t2,
                                                   const C2<float>* __restrict__ twist32,
                                                   const C2<double>* __restrict__ untw64,
                                                   const f16x8_t* __restrict__ fwd_a16,
                                                   const C2<float>* __restrict__ fwd_t2) {
    constexpr int M = N / 2, R1 = M / 64;
    __shared__ C2<float> s_buf[xbuf_elems<M>()];
    const int lane = threadIdx.x;
    const size_t g = blockIdx.x;
    DitTw<float> tw;
    load_dit_tw(tw, t2, t1, lane);
    C2<float> v[R1], tws[R1];
    if (inverse != 1) {
        const int32_t* p = (const int32_t*)in + g * N;
        if (inverse == 2) {
            FwdMfmaA am;
            load_fwd_mfma_a(am, fwd_a16, lane);
            uint32_t dw[R1];
#pragma unroll
            for (int aa = 0; aa < R1; ++aa) {
                // the coefficients of register aa as the bootstrap kernel reads them
                const int j = K32_BLAYOUT ? bl_lane(lane) + bl_off(aa) : lane + 64 * aa;
                dw[aa] = pack_f16x2((float)p[j], (float)p[j + M]);
                tws[aa] = fwd_t2[lane * R1 + aa];
            }
            fft512_fwd_mfma<K32_BLAYOUT>(v, dw, am, tws, tw, s_buf, lane);
        } else {
#pragma unroll
            for (int aa = 0; aa < R1; ++aa) {
                v[aa] = C2<float>{(float)p[lane + 64 * aa], (float)p[lane + 64 * aa + M]};
                tws[aa] = twist32[lane * R1 + aa];
            }
            fft512_dit<+1, true>(v, tw, (const C2<float>*)tws, s_buf, lane);
        }
        double* o = out + g * N;
#pragma unroll
        for (int s = 0; s < R1 / 8; ++s)
#pragma unroll
            for (int k3 = 0; k3 < 8; ++k3) {
                const int k = lane_order_k<R1>(lane, s, k3);
                o[2 * k] = v[s * 8 + k3].x;
                o[2 * k + 1] = v[s * 8 + k3].y;
            }
    } else {
        const double* F = (const double*)in + g * N;
#pragma unroll
        for (int s = 0; s < R1 / 8; ++s)
#pragma unroll
            for (int k3 = 0; k3 < 8; ++k3) {
                const int k = lane_order_k<R1>(lane, s, k3);
                v[s * 8 + k3] = C2<float>{(float)F[2 * k], (float)F[2 * k + 1]};
            }
        fft512_dit<-1, false>(v, tw, (const C2<float>*)nullptr, s_buf, lane);
        double* o = out + g * N;
#pragma unroll
        for (int aa = 0; aa < R1; ++aa) {
            C2<double> z = cmulc(C2<double>{(double)v[aa].x, (double)v[aa].y}, untw64[lane * R1 + aa]);
            o[lane + 64 * aa] = z.x;
            o[lane + 64 * aa + M] = z.y;
        }
    }
}

__global__ void __launch_bounds__(64) k_decrypt(int32_t* __restrict__ bits, const int32_t* __restrict__ s,
                                                const int32_t* __restrict__ key, int n) {
    const int g = blockIdx.x, lane = threadIdx.x;
    const int32_t* x = s + (size_t)g * (n + 1);
    uint32_t acc = 0;
    for (int i = lane; i < n; i += 64) acc += (uint32_t)x[i] * (uint32_t)key[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += (uint32_t)__shfl_xor((int)acc, off, 64);
    if (lane == 0) bits[g] = ((int32_t)((uint32_t)x[n] - acc) > 0) ? 1 : 0;
}

// Dense FMA chains (8 independent per lane) to measure the attainable VALU rate of one precision.
// The multiplier and addend are moved into vector registers first: an FP32 instruction that reads a
// scalar register issues at half the rate of one whose operands are all vector registers when two
// or more wavefronts share a SIMD, so uniform operands would measure the slower form.
template <typename T>
__global__ void __launch_bounds__(256) k_valu_probe(T* __restrict__ out, T a, T b, int iters) {
    T v[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) v[c] = (T)(threadIdx.x + c);
    T av = a, bv = b;
    asm volatile("" : "+v"(av), "+v"(bv));
    for (int it = 0; it < iters; ++it)
#pragma unroll
        for (int c = 0; c < 8; ++c) v[c] = v[c] * av + bv;
    T s = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) s += v[c];
    out[(size_t)blockIdx.x * 256 + threadIdx.x] = s;
}

// =============================================================================================
// host side: twiddle tables and launchers
// =============================================================================================

static long double pi_l() { return 3.14159265358979323846264338327950288L; }

// e^{i pi num / N} with the exponent reduced mod 2N before scaling
static void cis_pi_over_N(long long num, int N, double* re, double* im) {
    long long e = num % (2LL * N);
    if (e < 0) e += 2LL * N;
    long double ang = pi_l() * (long double)e / (long double)N;
    *re = (double)cosl(ang);
    *im = (double)sinl(ang);
}

hipError_t build_twiddles(DevTwiddles& tw, int N) {
    const int M = N / 2, R1 = M / 64;
    std::vector<double> t1f(2 * 64 * R1), t2(2 * 64 * 8), untw(2 * 64 * R1), wa(2 * R1);
    std::vector<float> t1p(2 * 64 * R1), t2f(2 * 64 * 8), t2d(2 * 64 * 8), tws(2 * 64 * R1);
    for (int L = 0; L < 64; ++L) {
        for (int k1 = 0; k1 < R1; ++k1) {
            double re, im;
            cis_pi_over_N((long long)L * (1 + 4 * k1), N, &re, &im);     // w^{L(1+4k1)}
            t1f[2 * (L * R1 + k1)] = re; t1f[2 * (L * R1 + k1) + 1] = im;
            cis_pi_over_N((long long)4 * L * k1, N, &re, &im);           // W_M^{L k1} = w^{4 L k1}
            t1p[2 * (L * R1 + k1)] = (float)re; t1p[2 * (L * R1 + k1) + 1] = (float)im;
            cis_pi_over_N((long long)64 * k1 + L, N, &re, &im);          // w^{64a + L}, a = k1
            tws[2 * (L * R1 + k1)] = (float)re; tws[2 * (L * R1 + k1) + 1] = (float)im;
            untw[2 * (L * R1 + k1)] = re / M; untw[2 * (L * R1 + k1) + 1] = im / M;
        }
        for (int k2 = 0; k2 < 8; ++k2) {
            // W_64^{c k2} = w^{(2N/64) c k2}
            double re, im;
            cis_pi_over_N((long long)(2 * N / 64) * (L & 7) * k2, N, &re, &im);
            t2[2 * (L * 8 + k2)] = re; t2[2 * (L * 8 + k2) + 1] = im;
            t2f[2 * (L * 8 + k2)] = (float)re; t2f[2 * (L * 8 + k2) + 1] = (float)im;
            // decimation-in-time pass-2 input twiddle W_64^{(L>>3) b}, b = k2
            cis_pi_over_N((long long)(2 * N / 64) * (L >> 3) * k2, N, &re, &im);
            t2d[2 * (L * 8 + k2)] = (float)re; t2d[2 * (L * 8 + k2) + 1] = (float)im;
        }
    }
    for (int a = 0; a < R1; ++a) {
        double re, im;
        cis_pi_over_N((long long)64 * a, N, &re, &im);
        wa[2 * a] = re; wa[2 * a + 1] = im;
    }
    // matrix-core pass 1 of the split forward transform (N = 1024, fft512_fwd_mfma): the real form of
    // G[k1][a] = W_8^{a k1} w^{64a}, rows 2 k1 + (0 re, 1 im), columns 2 a + (0 re, 1 im), split
    // into two FP16 pieces, and the pass-2 input twiddles W_64^{b k1} w^{8b + c} of lane c + 8 k1
    std::vector<_Float16> fa(64 * 16);
    std::vector<float> ft2(2 * 64 * 8);
    if (N == 1024) {
        double A[16][16];
        for (int k1 = 0; k1 < 8; ++k1)
            for (int a = 0; a < 8; ++a) {
                double gr, gi;
                cis_pi_over_N((long long)(2 * N / 8) * a * k1 + 64LL * a, N, &gr, &gi);
                A[2 * k1][2 * a] = gr; A[2 * k1][2 * a + 1] = -gi;
                A[2 * k1 + 1][2 * a] = gi; A[2 * k1 + 1][2 * a + 1] = gr;
            }
        // pieces P[0] = fp16(A), P[1] = fp16(A - P[0])
        _Float16 P[2][16][16];
        for (int m = 0; m < 16; ++m)
            for (int k = 0; k < 16; ++k) {
                P[0][m][k] = (_Float16)A[m][k];
                P[1][m][k] = (_Float16)(A[m][k] - (double)P[0][m][k]);
            }
        // lane L: entries 0-7 the two pieces in the K = 16 layout (A[L & 15][4 (L >> 4) + i]),
        // entries 8-15 the K = 32 layout with the pieces stacked along K (A[L & 15][8 (L >> 4) + i],
        // columns 16-31 being piece 2)
        for (int L = 0; L < 64; ++L) {
            for (int p = 0; p < 2; ++p)
                for (int i = 0; i < 4; ++i) fa[L * 16 + 4 * p + i] = P[p][L & 15][4 * (L >> 4) + i];
            for (int i = 0; i < 8; ++i) {
                const int k = 8 * (L >> 4) + i;
                fa[L * 16 + 8 + i] = P[k >> 4][L & 15][k & 15];
            }
        }
        for (int L = 0; L < 64; ++L)
            for (int b = 0; b < 8; ++b) {
                double re, im;
                cis_pi_over_N((long long)(2 * N / 64) * (L >> 3) * b + 8LL * b + (L & 7), N, &re, &im);
                ft2[2 * (L * 8 + b)] = (float)re; ft2[2 * (L * 8 + b) + 1] = (float)im;
            }
    }
    // two-wavefront form (N = 2048): pass-1 twiddles of the 512-point transform of wavefront v with
    // its lane twist w^{(1+4v) L} folded in, and its register twist constants w^{64 a (1+4v)}
    std::vector<double> t1w(2 * 2 * 64 * 8), rw(2 * 2 * 8), rws(2 * 2 * 8);
    std::vector<float> t1w32(t1w.size()), rw32(rw.size());
    if (N == 2048) {
        for (int v = 0; v < 2; ++v) {
            for (int L = 0; L < 64; ++L)
                for (int k1 = 0; k1 < 8; ++k1) {
                    double re, im;
                    cis_pi_over_N((long long)L * (1 + 4 * v + 8 * k1), N, &re, &im);
                    const size_t o = 2 * ((size_t)(v * 64 + L) * 8 + k1);
                    t1w[o] = re; t1w[o + 1] = im;
                    t1w32[o] = (float)re; t1w32[o + 1] = (float)im;
                }
            for (int a = 0; a < 8; ++a) {
                double re, im;
                cis_pi_over_N((long long)64 * a * (1 + 4 * v), N, &re, &im);
                rw[2 * (v * 8 + a)] = re; rw[2 * (v * 8 + a) + 1] = im;
                rw32[2 * (v * 8 + a)] = (float)re; rw32[2 * (v * 8 + a) + 1] = (float)im;
                rws[2 * (v * 8 + a)] = re / 1024; rws[2 * (v * 8 + a) + 1] = im / 1024;
            }
        }
    }
    // N = 2048 split path, matrix-core form of the forward (fft512_fwd2w_mfma): per parity v the
    // real form of X[k1] = sum_a W_8^{a k1} w^{64a(1+4v)} (P0_a + c_v P1_a), c_v = (-1)^v e^{i pi/4},
    // rows 2 k1 + (0 re, 1 im), columns 4a + (P0 re, P0 im, P1 re, P1 im), in two FP16 pieces, lane L
    // holding A_v[L & 15][8 (L >> 4) + i]; and the pass-1 twiddles w^{L(1+4v+8k1)} in pass 2's layout
    std::vector<_Float16> w2a(2 * 64 * 2 * 8);
    std::vector<float> w2t(2 * 2 * 64 * 8);
    if (N == 2048) {
        for (int v = 0; v < 2; ++v) {
            double A[16][32];
            for (int k1 = 0; k1 < 8; ++k1)
                for (int a = 0; a < 8; ++a) {
                    const long long e0 = (long long)(2 * N / 8) * a * k1 + 64LL * a * (1 + 4 * v);
                    double g0r, g0i, g1r, g1i;
                    cis_pi_over_N(e0, N, &g0r, &g0i);
                    cis_pi_over_N(e0 + N / 4 + (long long)N * v, N, &g1r, &g1i);   // times (-1)^v e^{i pi/4}
                    A[2 * k1][4 * a + 0] = g0r; A[2 * k1][4 * a + 1] = -g0i;
                    A[2 * k1][4 * a + 2] = g1r; A[2 * k1][4 * a + 3] = -g1i;
                    A[2 * k1 + 1][4 * a + 0] = g0i; A[2 * k1 + 1][4 * a + 1] = g0r;
                    A[2 * k1 + 1][4 * a + 2] = g1i; A[2 * k1 + 1][4 * a + 3] = g1r;
                }
            for (int L = 0; L < 64; ++L)
                for (int i = 0; i < 8; ++i) {
                    const double x = A[L & 15][8 * (L >> 4) + i];
                    const _Float16 h0 = (_Float16)x;
                    const _Float16 h1 = (_Float16)(x - (double)h0);
                    w2a[((v * 64 + L) * 2 + 0) * 8 + i] = h0;
                    w2a[((v * 64 + L) * 2 + 1) * 8 + i] = h1;
                }
            for (int L = 0; L < 64; ++L)
                for (int b = 0; b < 8; ++b) {
                    double re, im;
                    const long long Lv = 8 * b + (L & 7), k1 = L >> 3;
                    cis_pi_over_N(Lv * (1 + 4 * v + 8 * k1), N, &re, &im);
                    w2t[2 * ((v * 64 + L) * 8 + b)] = (float)re;
                    w2t[2 * ((v * 64 + L) * 8 + b) + 1] = (float)im;
                }
        }
    }
    // 4-bootstraps-per-wavefront form (N = 1024): pass-2 twiddles T[i][l + 16 s] =
    // e^{i pi i (4 (l + 16 s) + 1) / N}, table [l < 16][s < 2][i < 16]
    std::vector<float> q4(2 * 16 * 2 * 16);
    if (N == 1024) {
        for (int l = 0; l < 16; ++l)
            for (int s = 0; s < 2; ++s)
                for (int i = 0; i < 16; ++i) {
                    double re, im;
                    cis_pi_over_N((long long)i * (4 * (l + 16 * s) + 1), N, &re, &im);
                    const size_t o = 2 * ((size_t)(l * 2 + s) * 16 + i);
                    q4[o] = (float)re; q4[o + 1] = (float)im;
                }
    }
    hipError_t e;
    auto up = [&](void** dst, const void* src, size_t bytes) -> hipError_t {
        hipError_t r = hipMalloc(dst, bytes);
        if (r != hipSuccess) return r;
        return hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
    };
    tw.N = N;
    if ((e = up(&tw.t1_fold64, t1f.data(), t1f.size() * 8)) != hipSuccess) return e;
    if ((e = up(&tw.t2_64, t2.data(), t2.size() * 8)) != hipSuccess) return e;
    if ((e = up(&tw.t1_plain32, t1p.data(), t1p.size() * 4)) != hipSuccess) return e;
    if ((e = up(&tw.t2_32, t2f.data(), t2f.size() * 4)) != hipSuccess) return e;
    if ((e = up(&tw.t2dit32, t2d.data(), t2d.size() * 4)) != hipSuccess) return e;
    if ((e = up(&tw.twist32, tws.data(), tws.size() * 4)) != hipSuccess) return e;
    if ((e = up(&tw.untw64, untw.data(), untw.size() * 8)) != hipSuccess) return e;
    if ((e = up(&tw.wa64, wa.data(), wa.size() * 8)) != hipSuccess) return e;
    if (N == 1024 && (e = up(&tw.q4tw, q4.data(), q4.size() * 4)) != hipSuccess) return e;
    if (N == 1024 && (e = up(&tw.fwd_a16, fa.data(), fa.size() * 2)) != hipSuccess) return e;
    if (N == 1024 && (e = up(&tw.fwd_t2, ft2.data(), ft2.size() * 4)) != hipSuccess) return e;
    if (N == 2048) {
        if ((e = up(&tw.t1w64, t1w.data(), t1w.size() * 8)) != hipSuccess) return e;
        if ((e = up(&tw.t1w32, t1w32.data(), t1w32.size() * 4)) != hipSuccess) return e;
        if ((e = up(&tw.rw64, rw.data(), rw.size() * 8)) != hipSuccess) return e;
        if ((e = up(&tw.rw32, rw32.data(), rw32.size() * 4)) != hipSuccess) return e;
        if ((e = up(&tw.rws64, rws.data(), rws.size() * 8)) != hipSuccess) return e;
        if ((e = up(&tw.w2_a16, w2a.data(), w2a.size() * 2)) != hipSuccess) return e;
        if ((e = up(&tw.w2_t1p32, w2t.data(), w2t.size() * 4)) != hipSuccess) return e;
    }
    return hipSuccess;
}

void free_twiddles(DevTwiddles& tw) {
    void* ps[] = {tw.t1_fold64, tw.t2_64, tw.t1_plain32, tw.t2_32, tw.t2dit32, tw.twist32, tw.untw64, tw.wa64,
                  tw.t1w64, tw.t1w32, tw.rw64, tw.rw32, tw.rws64, tw.q4tw, tw.fwd_a16, tw.fwd_t2, tw.w2_a16, tw.w2_t1p32};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    tw = DevTwiddles();
}

bool params_supported(const DevParams& p) {
    const bool ring = (p.N == 1024 && p.l == 3 && p.Bgbit == 7) || (p.N == 2048 && p.l == 3 && p.Bgbit == 8);
    return ring && p.n > 0 && p.n <= kMaxLweN && p.ks_basebit == 2 && (p.ks_t == 8 || p.ks_t == 10);
}

// ---------------------------------------------------------------------------------------------
// N = 2048: each folded transform (M = 1024 points) is split over the two wavefronts of a
// bootstrap, so that every wavefront keeps the 8-points-per-lane register shape of N = 1024.
// With z_j = P_j w^j (P_j = p_j + i p_{j+1024}, w = e^{i pi/2048}) and the split j = j' + 512h of
// the transform sum, the frequencies of parity v are a 512-point transform of
//   y_v[j'] = w^{(1+4v) j'} (P_{j'} + (-1)^v e^{i pi/4} P_{j'+512}),        j' < 512,
// because w^512 = e^{i pi/4} and the odd half carries W_1024^{j'} = w^{4j'}.  Wavefront v builds
// y_v from all 2048 coefficients (no exchange on the way in) and runs the 512-point transform with
// the lane part of its twist folded into the pass-1 twiddles; its result is F_{2m+v}, m in lane
// order.  The key is stored in the same [parity][register][lane] order.  On the way back each
// wavefront inverts its own half; with E = wavefront 0's and O = wavefront 1's output after the
// register untwist,
//   (p_{j'}, p_{j'+1024})       = (E + O) / 1024                     (wavefront 0 keeps these)
//   (p_{j'+512}, p_{j'+1536})   = (E - O) e^{-i pi/4} / 1024          (wavefront 1 keeps these)
// which needs one exchange of the inverse outputs through LDS.
// ---------------------------------------------------------------------------------------------
constexpr int W2_N = 2048, W2_M = 1024, W2_H = 512;
// offset (in C2<float>) of the upper half of a split-path wavefront buffer: past the transform
// exchange area, 4 KB of C2<float>[512] to the buffer's end
constexpr int W2_XHALF = 576;

template <typename T>
__device__ __forceinline__ C2<T> w2_c() {
    const T h = (T)0.70710678118654752440;   // e^{i pi/4}
    return C2<T>{h, h};
}

// y_v[j'] before its lane twist: register twist w^{64a(1+4v)} (1 for a = 0) times P0 +- c P1
template <typename T>
__device__ __forceinline__ C2<T> w2_prep(C2<T> p0, C2<T> p1, C2<T> rw, int wave, bool first) {
    const C2<T> t = cmul(p1, w2_c<T>());
    const C2<T> y = wave ? C2<T>{p0.x - t.x, p0.y - t.y} : C2<T>{p0.x + t.x, p0.y + t.y};
    return first ? y : cmul(y, rw);
}

// the pair this wavefront keeps from its own inverse output X (register-twisted, unscaled) and the
// partner's output Y, both already untwisted: wavefront 0 -> (E + O)/1024, 1 -> (E - O) c^* /1024
__device__ __forceinline__ C2<double> w2_pair(C2<double> own, C2<double> other, int wave) {
    const double s = 1.0 / (double)W2_M;
    if (!wave) return C2<double>{(own.x + other.x) * s, (own.y + other.y) * s};
    const C2<double> d{(other.x - own.x) * s, (other.y - own.y) * s};   // E - O, E is the partner's
    return cmulc(d, w2_c<double>());
}

// Split path: y_v[j'] before its lane twist, (P0 + hc (p1.x - p1.y, p1.x + p1.y)) w^{64a(1+4v)},
// hc = +-cos(pi/4) the sign of wavefront v: P0 +- c P1 in two subtractions and two FMAs, then the
// register twist (skipped for a = 0)
__device__ __forceinline__ C2<float> w2_prep32(float p0x, float p0y, float p1x, float p1y, float hc, C2<float> rw,
                                               bool first) {
    const C2<float> y{__builtin_fmaf(hc, p1x - p1y, p0x), __builtin_fmaf(hc, p1x + p1y, p0y)};
    return first ? y : cmul(y, rw);
}

// Split path: the pair this wavefront keeps from its own and the partner's inverse outputs, each
// untwisted and scaled by 1/1024 by its own wavefront: wavefront 0 E + O, wavefront 1 (E - O) c^*
__device__ __forceinline__ C2<double> w2_pair_own(C2<double> own, C2<double> other, int wave) {
    if (!wave) return C2<double>{own.x + other.x, own.y + other.y};
    const double h = 0.70710678118654752440;
    const double dx = other.x - own.x, dy = other.y - own.y;
    return C2<double>{h * (dx + dy), h * (dy - dx)};
}

// coefficient index (within one polynomial) of the real / imaginary part that wavefront `wave`
// keeps for register a
__device__ __forceinline__ int w2_coef(int wave, int lane, int a, int imag) {
    return lane + 64 * a + 512 * wave + 1024 * imag;
}

// Touches every 128-byte line of this wavefront's half of a key row (2 polynomials x 512 entries of
// 16 bytes): one dword per lane per polynomial, folded into `sink` so the loads stay live.  Issued
// after a row's key loads, it brings the next row into L2 while the next transform runs; that row's
// own loads then hit L2 instead of stalling on the ~2% of lines that would come from memory.
__device__ __forceinline__ void w2_warm_row(const void* row_half, int lane, float& sink) {
    const float* p = (const float*)row_half;
    sink += p[lane * 32] + p[W2_M * 4 + lane * 32];
}

// The accumulators of the N = 2048 kernels carry mirrors of their ends in LDS: the MW words below
// p0 repeat p0's last MW coefficients and the MW words past p1 repeat p1's first, so that the
// rotation reads every coefficient at a plain offset from a non-negative base (w2_rotate).  The
// rotation reads registers in groups of G, so MW = 64 (G - 1): the split kernel uses G = 8 (fewer
// base registers), the FP64 kernel G = 4 (its workgroup has LDS for 192-word mirrors only).
template <int G>
__host__ __device__ constexpr int w2_mirror() { return 64 * (G - 1); }

// Lane L copies k = L, L + 64, ...: sources N - MW + k and N + k are coefficients that lane L
// itself wrote in the prologue (both offsets are multiples of 64), so program order suffices.
template <int G>
__device__ __forceinline__ void w2_mirror_fill(int32_t* s_acc, int lane) {
    constexpr int MW = w2_mirror<G>();
    static_assert((W2_N - MW) % 64 == 0 && W2_N % 64 == 0, "mirror sources in the copying lane");
    for (int k = lane; k < MW; k += 64) {
        s_acc[k - MW] = s_acc[W2_N - MW + k];
        s_acc[2 * W2_N + k] = s_acc[W2_N + k];
    }
}

// After wavefront `wave` has written n0 = p_q[w2_coef(wave, lane, aa, 0)] and n1 (its imaginary
// partner): p0's coefficients from N - MW up are wavefront 1's imaginary parts with
// aa >= 8 - MW / 64, and p1's below MW are wavefront 0's real parts with aa < MW / 64.
template <int G>
__device__ __forceinline__ void w2_mirror_update(int32_t* s_acc, int q, int wave, int aa, int j0, int j1, int32_t n0,
                                                 int32_t n1) {
    constexpr int MW = w2_mirror<G>();
    if (q == 0 && wave == 1 && aa >= 8 - MW / 64) s_acc[j1 - W2_N] = n1;
    if (q == 1 && wave == 0 && aa < MW / 64) s_acc[j0 + W2_N] = n0;
}

// u[r] = ((X^a - 1) p_q + offset) ^ offset at coefficient j = lane + 64 r (r < 32), or
// (p_q + offset) ^ offset when !ROT.  The rotation reads p at e = (j - a) mod 2N, negated when
// e >= N.  With d = lane - a and E = (d mod 2N) >> 6, e >= N exactly when bit 5 of (E + r) mod 64
// is set, so bit r of the low word of rotr64(0xFFFFFFFF00000000, E) is the sign of register r (one
// v_alignbit).  Registers r = G g + rr read p[(b_g + 64 rr) mod N], b_g = (d + 64 G g) mod N.  p1
// is read at b_g + 64 rr < N + MW, which the mirror past it covers; p0 from the mirror below it at
// c_g + 64 rr - MW, c_g = (b_g + MW) mod N, in [-MW, N).  Every base register is non-negative (the
// MW goes into the loads' constant offsets: a negative base register with a positive offset gave
// wrong results) and no read needs wrap arithmetic.  +-p_e = (p_e + m) ^ m for m = 0 or -1, and the
// opaque offset term keeps the compiler from reassociating the sum, so that the xor and the add
// fuse into one v_xad_u32: four instructions per coefficient and one sign extract per register.
template <int G, bool ROT, uint32_t OFF, bool XOFF = true>
__device__ __forceinline__ void w2_rotate(uint32_t* u, const int32_t* s_acc, int q, int a, int lane) {
    constexpr int N = W2_N, NC = 32, MW = w2_mirror<G>(), NG = NC / G;
    const int32_t* pq = s_acc + q * N;
    if constexpr (ROT) {
        const int d = lane - a;
        const uint32_t E = (uint32_t)(d >> 6) & 63u;
        const uint32_t hw = (E & 32u) ? 0u : 0xFFFFFFFFu;
        const uint32_t sgn = __builtin_amdgcn_alignbit(hw, ~hw, E & 31u);
        const int sh = q ? 0 : MW;
        int base[NG];
#pragma unroll
        for (int g = 0; g < NG; ++g) base[g] = (d + 64 * G * g + sh) & (N - 1);
#pragma unroll
        for (int r = 0; r < NC; ++r) {
            const uint32_t pj = (uint32_t)pq[lane + 64 * r];
            const uint32_t sv = (uint32_t)pq[base[r / G] + 64 * (r % G) - sh];
            const uint32_t m = (uint32_t)((int32_t)(sgn << (31 - r)) >> 31);
            uint32_t c = OFF - pj;
            asm("" : "+v"(c));
            u[r] = ((sv + m) ^ m) + c;
            if constexpr (XOFF) u[r] ^= OFF;
        }
    } else {
#pragma unroll
        for (int r = 0; r < NC; ++r) u[r] = ((uint32_t)pq[lane + 64 * r] + OFF) ^ (XOFF ? OFF : 0u);
    }
}

// mirror group sizes of the N = 2048 kernels (see w2_mirror): the split latency form, the FP64
// kernel, and the split two-wavefront kernel (w2_rotate_b's 128-word mirrors, G = 3, or w2_rotate's)
constexpr int W2_G_SPLIT = 8, W2_G_FP64 = 4, W2_G_SPLIT2W = TFHE_W2_BLAYOUT ? 3 : W2_G_SPLIT;

// w2_rotate for the operand layout (bl_off): u[r] = ((X^a - 1) p_q + offset) ^ offset at
// coefficient j = bl_lane(lane) + bl_off(r).  With d = j0 - a (j0 = bl_lane(lane)), register r reads p at
// e = (d + bl_off(r)) mod 2N, negated when bit 11 of d + bl_off(r) is set.  Quarter k = r >> 3
// reads from one base b_k = (d + 512 k) mod N at offsets bl_off(r & 7) <= 112, so 128-word mirrors
// (w2_mirror<3>) cover every read, p0's through the shifted base as in w2_rotate.  The reads of
// one register put the four rows of 16 lanes 128 words apart.
template <int MW, bool ROT, uint32_t OFF, bool XOFF = true>
__device__ __forceinline__ void w2_rotate_b(uint32_t* u, const int32_t* s_acc, int q, int a, int lane) {
    constexpr int N = W2_N, NC = 32;
    static_assert(MW >= 112, "mirrors cover the largest in-quarter offset");
    const int32_t* pq = s_acc + q * N;
    const int j0 = bl_lane(lane);
    if constexpr (ROT) {
        const int d = j0 - a;
        const uint32_t x = (uint32_t)d << 20;   // bit 11 of d + offset -> bit 31
        const int sh = q ? 0 : MW;
        int base[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) base[k] = (d + 512 * k + sh) & (N - 1);
#pragma unroll
        for (int r = 0; r < NC; ++r) {
            const uint32_t pj = (uint32_t)pq[j0 + bl_off(r)];
            const uint32_t sv = (uint32_t)pq[base[r >> 3] + bl_off(r & 7) - sh];
            const uint32_t m = (uint32_t)((int32_t)(x + ((uint32_t)bl_off(r) << 20)) >> 31);
            uint32_t c = OFF - pj;
            asm("" : "+v"(c));
            u[r] = ((sv + m) ^ m) + c;
            if constexpr (XOFF) u[r] ^= OFF;
        }
    } else {
#pragma unroll
        for (int r = 0; r < NC; ++r) u[r] = ((uint32_t)pq[j0 + bl_off(r)] + OFF) ^ (XOFF ? OFF : 0u);
    }
}

// Split-FP32 step at N = 2048, one wavefront's share.  ROT as in step32.  Each wavefront untwists
// its own inverse outputs in FP64 (rws: register constants with the 1/1024) and the untwisted values
// are exchanged, so no wavefront untwists its partner's.  sx_own / sx_other: 9 KB per wavefront,
// the transform exchange buffer (C2<float>) and two halves for the output exchange.
template <int L, int BG, bool ROT>
__device__ __forceinline__ void step2w32(int32_t* s_acc, void* sx_own, const void* sx_other, int a,
                                         const float4* __restrict__ kr, const LaneTw<float, 8>& tw, const C2<float>* __restrict__ rw32,
                                         const C2<double>* __restrict__ rws, const f16x8_t* a2, int wave, int lane,
                                         int split_bits, double* raw, double* raw_hi, bool work) {
    constexpr int N = W2_N, M = W2_M, R1 = 8, NC = 32;
    constexpr uint32_t OFF = decomp_offset<L, BG>();
    C2<float>* sbuf = (C2<float>*)sx_own;
    const float hc = wave ? -0.70710678118654752440f : 0.70710678118654752440f;
    C2<float> h0[R1], l0[R1], h1[R1], l1[R1];
#pragma unroll
    for (int r = 0; r < R1; ++r) {
        h0[r] = C2<float>{0.f, 0.f}; l0[r] = C2<float>{0.f, 0.f};
        h1[r] = C2<float>{0.f, 0.f}; l1[r] = C2<float>{0.f, 0.f};
    }
    // an idle bootstrap (zero rotation, or past the batch end) skips the arithmetic but still takes
    // part in every workgroup barrier below
    if (work) {
    uint32_t u[NC];
#if TFHE_W2_BLAYOUT
    auto rotate = [&](int q) { w2_rotate_b<w2_mirror<W2_G_SPLIT2W>(), ROT, OFF, !DIGIT_PERM>(u, s_acc, q, a, lane); };
#else
    auto rotate = [&](int q) { w2_rotate<W2_G_SPLIT2W, ROT, OFF>(u, s_acc, q, a, lane); };
#endif
#if TFHE_W2_FWD_MFMA
    // the row's digits as FP16 pairs P0_a = (p_{j'}, p_{j'+1024}), P1_a = (p_{j'+512}, p_{j'+1536})
    (void)hc;
    (void)rw32;
    auto prep = [&](uint32_t* x0, uint32_t* x1, int lvl) {
#pragma unroll
        for (int aa = 0; aa < R1; ++aa) {
#if TFHE_W2_BLAYOUT && DIGIT_PERM
            static_assert(BG == 8, "byte-aligned digits");
            x0[aa] = digits8_f16x2(u[aa], u[aa + 16], lvl);
            x1[aa] = digits8_f16x2(u[aa + 8], u[aa + 24], lvl);
#else
            x0[aa] = pack_f16x2((float)digit_sx<BG>(u[aa], lvl), (float)digit_sx<BG>(u[aa + 16], lvl));
            x1[aa] = pack_f16x2((float)digit_sx<BG>(u[aa + 8], lvl), (float)digit_sx<BG>(u[aa + 24], lvl));
#endif
        }
    };
#else
    auto prep = [&](C2<float>* x, int lvl) {
#pragma unroll
        for (int aa = 0; aa < R1; ++aa)
            x[aa] = w2_prep32((float)digit_sx<BG>(u[aa], lvl), (float)digit_sx<BG>(u[aa + 16], lvl),
                              (float)digit_sx<BG>(u[aa + 8], lvl), (float)digit_sx<BG>(u[aa + 24], lvl), hc,
                              rw32[aa], aa == 0);
    };
#endif
    // rows t = (q, lvl) in order, software-pipelined: the row's first key polynomial is requested
    // before its transform, the second after it, and the next row's digits are prepared (rotating
    // the next polynomial first when q changes) while the second polynomial is in flight
    C2<float> v[R1];
#if TFHE_W2_FWD_MFMA
    uint32_t x0[R1], x1[R1];
#endif
    float4 k0[R1];
    // row 0's first key polynomial is requested before the rotation, which then hides it too
#pragma unroll
    for (int r = 0; r < R1; ++r) k0[r] = kr[wave * W2_H + lane + r * 64];
    rotate(0);
#if TFHE_W2_FWD_MFMA
    prep(x0, x1, 0);
#else
    prep(v, 0);
#endif
#pragma unroll
    for (int row = 0; row < 2 * L; ++row) {
        const float4* k = kr + (size_t)(row * 2) * M + wave * W2_H + lane;
        if (row > 0) {
#pragma unroll
            for (int r = 0; r < R1; ++r) k0[r] = k[r * 64];
        }
#if TFHE_W2_FWD_MFMA
        fft512_fwd2w_mfma<TFHE_W2_BLAYOUT>(v, x0, x1, a2, tw.t1, tw.t2, sbuf, lane);
#else
        fft_fwd<R1>(v, tw, sbuf, lane);
#endif
#pragma unroll
        for (int r = 0; r < R1; ++r) {
            cmac(h0[r], v[r], C2<float>{k0[r].x, k0[r].y});
            cmac(l0[r], v[r], C2<float>{k0[r].z, k0[r].w});
        }
        float4 k1[R1];
#pragma unroll
        for (int r = 0; r < R1; ++r) k1[r] = k[M + r * 64];
#if TFHE_W2_FWD_MFMA
        if (row + 1 < 2 * L) {
            if ((row + 1) % L == 0) rotate((row + 1) / L);
            prep(x0, x1, (row + 1) % L);
        }
#else
        C2<float> vn[R1];
        if (row + 1 < 2 * L) {
            if ((row + 1) % L == 0) rotate((row + 1) / L);
            prep(vn, (row + 1) % L);
        }
#endif
#pragma unroll
        for (int r = 0; r < R1; ++r) {
            cmac(h1[r], v[r], C2<float>{k1[r].x, k1[r].y});
            cmac(l1[r], v[r], C2<float>{k1[r].z, k1[r].w});
        }
        // a scheduling barrier between rows.  The next row's warm-up loads that used to sit here
        // cost 4.5% once the key requests moved ahead of the transforms; without the barrier the
        // step is 6% slower.  (The FP64 step keeps its warm-up loads: 30% slower without.)
        __builtin_amdgcn_sched_barrier(0);
#if !TFHE_W2_FWD_MFMA
        if (row + 1 < 2 * L) {
#pragma unroll
            for (int r = 0; r < R1; ++r) v[r] = vn[r];
        }
#endif
    }
#if TFHE_W2_FWD_MFMA
    fft512_inv2w_t1p(h0, tw.t1, tw.t2, sbuf, lane);
    fft512_inv2w_t1p(l0, tw.t1, tw.t2, sbuf, lane);
    fft512_inv2w_t1p(h1, tw.t1, tw.t2, sbuf, lane);
    fft512_inv2w_t1p(l1, tw.t1, tw.t2, sbuf, lane);
#else
    fft_inv<R1>(h0, tw, sbuf, lane);
    fft_inv<R1>(l0, tw, sbuf, lane);
    fft_inv<R1>(h1, tw, sbuf, lane);
    fft_inv<R1>(l1, tw, sbuf, lane);
#endif
    }
    const double hs = ldexp(1.0, split_bits);
    // one output polynomial at a time: exchange hi, combine and round it, then lo.  The
    // exchanged values are rounded to FP32 and alternate between two 4 KB halves of the buffer, so
    // a round needs one barrier: a half is written again two rounds later, after the partner has
    // passed the barrier that follows its read.
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        uint32_t rh[2 * R1];
        C2<double> ph[R1];
#pragma unroll
        for (int part = 0; part < 2; ++part) {
            const C2<float>* X = part ? (q ? l1 : l0) : (q ? h1 : h0);
            C2<float>* xo = (C2<float>*)sx_own + part * W2_XHALF;
            const C2<float>* xp = (const C2<float>*)sx_other + part * W2_XHALF;
            C2<double> own[R1];
            if (work) {
#pragma unroll
                for (int aa = 0; aa < R1; ++aa) {
                    own[aa] = cmulc(C2<double>{(double)X[aa].x, (double)X[aa].y}, rws[aa]);
                    xo[aa * 64 + lane] = C2<float>{(float)own[aa].x, (float)own[aa].y};
                }
            }
            __syncthreads();
#pragma unroll
            for (int aa = 0; aa < R1; ++aa) {
                if (!work) break;
                const C2<float> o = xp[aa * 64 + lane];
                const C2<double> z = w2_pair_own(own[aa], C2<double>{(double)o.x, (double)o.y}, wave);
                if (!part) {
                    rh[2 * aa] = round_mod32(z.x);
                    rh[2 * aa + 1] = round_mod32(z.y);
                    ph[aa] = z;
                    if (raw_hi) {   // the hi product before rounding, in units of 2^split_bits
                        raw_hi[q * N + w2_coef(wave, lane, aa, 0)] = z.x;
                        raw_hi[q * N + w2_coef(wave, lane, aa, 1)] = z.y;
                    }
                } else {
                    const uint32_t c0 = (rh[2 * aa] << split_bits) + round_mod32(z.x);
                    const uint32_t c1 = (rh[2 * aa + 1] << split_bits) + round_mod32(z.y);
                    const int j0 = q * N + w2_coef(wave, lane, aa, 0), j1 = q * N + w2_coef(wave, lane, aa, 1);
                    if (raw) { raw[j0] = ph[aa].x * hs + z.x; raw[j1] = ph[aa].y * hs + z.y; }
                    if (ROT) {
                        const int32_t n0 = (int32_t)((uint32_t)s_acc[j0] + c0);
                        const int32_t n1 = (int32_t)((uint32_t)s_acc[j1] + c1);
                        s_acc[j0] = n0;
                        s_acc[j1] = n1;
                        w2_mirror_update<W2_G_SPLIT2W>(s_acc, q, wave, aa, j0, j1, n0, n1);
                    } else {
                        s_acc[j0] = (int32_t)c0;
                        s_acc[j1] = (int32_t)c1;
                    }
                }
            }
        }
    }
    // the partner reads the upper half after the last barrier above, and every accumulator
    // coefficient is read at the start of the next step
    __syncthreads();
}

// FP64 step at N = 2048, one wavefront's share.
template <int L, int BG, bool ROT>
__device__ __forceinline__ void step2w64(int32_t* s_acc, C2<double>* sx_own, const C2<double>* sx_other, int a,
                                         const C2<double>* __restrict__ kr, const C2<double>* __restrict__ kr_next,
                                         const LaneTwS<double, 8>& tw0,
                                         const C2<double>* __restrict__ rw_own, const C2<double>* __restrict__ rw_other,
                                         int wave, int lane, double* raw, bool work, float& sink) {
    constexpr int N = W2_N, M = W2_M, R1 = 8, NC = 32;
    constexpr uint32_t OFF = decomp_offset<L, BG>();
    // an offset the compiler cannot see through keeps the pass-2 twiddle reads at their uses (held
    // across steps they would cost 28 registers).  The statement is not volatile, so it is no memory
    // clobber and the uniform constant loads after it can be scalar loads (0.5% per step); its input
    // a changes from step to step, so it is not hoisted out of the step loop.
    int t2o = 0;
    asm("" : "+v"(t2o) : "s"(a));
    LaneTwS<double, 8> tw = tw0;
    tw.t2 = tw0.t2 + t2o;
    C2<double> o0[R1], o1[R1];
#pragma unroll
    for (int r = 0; r < R1; ++r) { o0[r] = C2<double>{0.0, 0.0}; o1[r] = C2<double>{0.0, 0.0}; }
    if (work) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        uint32_t u[NC];
        w2_rotate<W2_G_FP64, ROT, OFF>(u, s_acc, q, a, lane);
#pragma unroll
        for (int lvl = 0; lvl < L; ++lvl) {
            const int row = q * L + lvl;
            C2<double> v[R1];
#pragma unroll
            for (int aa = 0; aa < R1; ++aa) {
                const C2<double> p0{(double)digit_sx<BG>(u[aa], lvl), (double)digit_sx<BG>(u[aa + 16], lvl)};
                const C2<double> p1{(double)digit_sx<BG>(u[aa + 8], lvl), (double)digit_sx<BG>(u[aa + 24], lvl)};
                v[aa] = w2_prep(p0, p1, rw_own[aa], wave, aa == 0);
            }
            const C2<double>* k = kr + (size_t)((q * L + lvl) * 2) * M + wave * W2_H + lane;
            // the row's first EK key values are requested before the transform, whose length hides
            // their latency (more spill at 256 VGPRs); the rest after it
            constexpr int EK = 3;
            C2<double> k0[EK];
#pragma unroll
            for (int r = 0; r < EK; ++r) k0[r] = k[r * 64];
            fft_fwd<R1, double, false, false>(v, tw, sx_own, lane);
#pragma unroll
            for (int r = 0; r < R1; ++r) cmac(o0[r], v[r], r < EK ? k0[r < EK ? r : 0] : k[r * 64]);
#pragma unroll
            for (int r = 0; r < R1; ++r) cmac(o1[r], v[r], k[M + r * 64]);
            __builtin_amdgcn_sched_barrier(0);
            w2_warm_row(row + 1 < 2 * L ? kr + (size_t)((row + 1) * 2) * M + wave * W2_H : kr_next + wave * W2_H, lane,
                        sink);
        }
    }
    fft_inv<R1, double, false, false>(o0, tw, sx_own, lane);
    fft_inv<R1, double, false, false>(o1, tw, sx_own, lane);
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const C2<double>* X = q ? o1 : o0;
        if (work) {
#pragma unroll
            for (int aa = 0; aa < R1; ++aa) sx_own[aa * 64 + lane] = X[aa];
        }
        __syncthreads();
#pragma unroll
        for (int aa = 0; aa < R1; ++aa) {
            if (!work) break;
            const C2<double> y = sx_other[aa * 64 + lane];
            const C2<double> own = aa ? cmulc(X[aa], rw_own[aa]) : X[aa];
            const C2<double> oth = aa ? cmulc(y, rw_other[aa]) : y;
            const C2<double> z = w2_pair(own, oth, wave);
            const int j0 = q * N + w2_coef(wave, lane, aa, 0), j1 = q * N + w2_coef(wave, lane, aa, 1);
            if (raw) { raw[j0] = z.x; raw[j1] = z.y; }
            const uint32_t c0 = round_mod32(z.x), c1 = round_mod32(z.y);
            if (ROT) {
                const int32_t n0 = (int32_t)((uint32_t)s_acc[j0] + c0);
                const int32_t n1 = (int32_t)((uint32_t)s_acc[j1] + c1);
                s_acc[j0] = n0;
                s_acc[j1] = n1;
                w2_mirror_update<W2_G_FP64>(s_acc, q, wave, aa, j0, j1, n0, n1);
            } else {
                s_acc[j0] = (int32_t)c0;
                s_acc[j1] = (int32_t)c1;
            }
        }
        // the partner's exchange buffer is read above and written again for the next polynomial,
        // and both wavefronts read every accumulator coefficient at the start of the next step
        __syncthreads();
    }
}

// Bootstrap at N = 2048, two wavefronts per bootstrap, GB bootstraps per workgroup (their key reads
// of a row come close together and share the CU's L1; the exchange barriers keep them in step).
// P = 0: FP64, P = 1: split FP32.
template <int L, int BG, int P, int GB>
__global__ void __launch_bounds__(128 * GB) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_boot2w(BootKArgs A) {
    constexpr int N = W2_N, M = W2_M, KPL = 2 * L;
    constexpr int XB = xbuf_elems<W2_H>();
    constexpr int T2S = 9;   // row stride of the LDS pass-2 twiddles: the 8 rows start on distinct banks
    // the accumulators carry mirrors of their ends (w2_rotate)
    constexpr int AM = w2_mirror<P ? W2_G_SPLIT2W : W2_G_FP64>();
    __shared__ int32_t s_acc_all[GB][2 * N + 2 * AM];
    // per wavefront C2<double>[XB]: FP64 transform exchanges; split-FP32 transform exchanges
    // (C2<float>[XB]) and its untwisted outputs, rounded to FP32, in two alternating halves
    __shared__ double s_xd[GB][2][2 * XB];
    __shared__ int16_t s_bara_all[GB][kMaxLweN];
    __shared__ C2<double> s_t2[P ? 1 : 8 * T2S];
    const int lane = threadIdx.x & 63;
    const int wid = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int b = wid >> 1, wave = wid & 1;
    const int g = blockIdx.x * GB + b;
    const bool active = g < A.L.count;
    int32_t* s_acc = s_acc_all[b] + AM;
    int16_t* s_bara = s_bara_all[b];
    if (active && wave == 0) {
        boot_prologue<N>(A, g, lane, s_acc, s_bara);
        // no fence: each lane copies only coefficients it wrote itself (w2_mirror_fill), and a
        // fence here would count as a possible write to global memory, turning the step's uniform
        // constant loads into vector loads
        w2_mirror_fill<P ? W2_G_SPLIT2W : W2_G_FP64>(s_acc, lane);
    }
    if constexpr (P == 0) {
        // lane c of the pass-2 table holds W_64^{c k2}, k2 < 8
        if (threadIdx.x < 64) s_t2[(threadIdx.x >> 3) * T2S + (threadIdx.x & 7)] = ((const C2<double>*)A.t2)[threadIdx.x];
    }
    __syncthreads();
    const C2<double>* rw64 = (const C2<double>*)A.t2w_rw64;
    const C2<double>* rw_own = rw64 + wave * 8;
    const C2<double>* rw_other = rw64 + (1 - wave) * 8;
    constexpr size_t ROWSET = (size_t)KPL * 2 * M;
    const size_t rs = (A.L.mode == 2 && active) ? (size_t)A.L.row_set[g] * ROWSET : 0;
    if constexpr (P == 1) {
        LaneTw<float, 8> tw;
        f16x8_t a2[2];
#if TFHE_W2_FWD_MFMA
        // t1: the pass-1 twiddles in pass 2's layout, and the matrix-core pieces of this parity
        load_lane_tw(tw, (const C2<float>*)A.fwd_t2 + wave * 64 * 8, (const C2<float>*)A.t2, lane);
        a2[0] = ((const f16x8_t*)A.fwd_a16)[(wave * 64 + lane) * 2];
        a2[1] = ((const f16x8_t*)A.fwd_a16)[(wave * 64 + lane) * 2 + 1];
#else
        load_lane_tw(tw, (const C2<float>*)A.t1 + wave * 64 * 8, (const C2<float>*)A.t2, lane);
#endif
        const C2<float>* rw32 = (const C2<float>*)A.t2w_rw32 + wave * 8;
        const C2<double>* rws = (const C2<double>*)A.untw64 + wave * 8;
        void* sx_own = s_xd[b][wave];
        const void* sx_other = s_xd[b][1 - wave];
        const float4* key = (const float4*)A.bk;
        if (A.L.mode == 2) {
            double* raw = (A.L.raw_out && active) ? A.L.raw_out + (size_t)g * 2 * N : nullptr;
            double* raw_hi = (A.L.raw_hi && active) ? A.L.raw_hi + (size_t)g * 2 * N : nullptr;
            step2w32<L, BG, false>(s_acc, sx_own, sx_other, 0, key + rs, tw, rw32, rws, a2, wave, lane,
                                   A.split_bits, raw, raw_hi, active);
        } else {
            for (int i = 0; i < A.n; ++i) {
                // a zero rotation skips the step (tfhe_blindRotate_FFT.cpp:114); a is the same in both
                // wavefronts of a bootstrap, and an idle bootstrap still meets every barrier
                const int a = active ? __builtin_amdgcn_readfirstlane((int)s_bara[i]) : 0;
                step2w32<L, BG, true>(s_acc, sx_own, sx_other, a, key + (size_t)i * ROWSET, tw, rw32, rws, a2, wave, lane,
                                      A.split_bits, nullptr, nullptr, a != 0);
            }
        }
    } else {
        LaneTwS<double, 8> tw;
#pragma unroll
        for (int r = 0; r < 8; ++r) tw.t1[r] = ((const C2<double>*)A.t1)[(wave * 64 + lane) * 8 + r];
        tw.t2 = s_t2 + (lane & 7) * T2S;
        C2<double>* sx_own = (C2<double>*)s_xd[b][wave];
        const C2<double>* sx_other = (const C2<double>*)s_xd[b][1 - wave];
        const C2<double>* key = (const C2<double>*)A.bk;
        // the warm-up loads are folded into sink, which is consumed once after the last step (an
        // empty asm statement inside the step would stop uniform loads from becoming scalar loads)
        float sink = 0.f;
        if (A.L.mode == 2) {
            double* raw = (A.L.raw_out && active) ? A.L.raw_out + (size_t)g * 2 * N : nullptr;
            step2w64<L, BG, false>(s_acc, sx_own, sx_other, 0, key + rs, key + rs, tw, rw_own, rw_other, wave, lane,
                                   raw, active, sink);
        } else {
            for (int i = 0; i < A.n; ++i) {
                const int a = active ? __builtin_amdgcn_readfirstlane((int)s_bara[i]) : 0;
                step2w64<L, BG, true>(s_acc, sx_own, sx_other, a, key + (size_t)i * ROWSET,
                                      key + (size_t)(i + 1 < A.n ? i + 1 : i) * ROWSET, tw, rw_own, rw_other, wave,
                                      lane, nullptr, a != 0, sink);
            }
        }
        asm volatile("" ::"v"(sink));
    }
    __syncthreads();
    if (active && wave == 0) boot_epilogue<N>(A, g, lane, s_acc);
}

// ---------------------------------------------------------------------------------------------
// N = 2048 latency form: one bootstrap per workgroup.  Wavefront w < 2l is (level lvl = w / 2,
// parity v = w % 2): for q = 0, 1 it rotates polynomial q, transforms its level-lvl digits into the
// parity-v half of the spectrum (the two-wavefront split above) and multiplies them by the
// parity-v half of key row (q, lvl), accumulating one product per output (FP64: the two output
// polynomials; split: each polynomial times the hi and the lo key half).  After a barrier,
// wavefront w < 2K takes parity v = w % 2 of output k = w / 2: it sums the l partials in a fixed
// order, inverts them, untwists its own half and leaves it in LDS (over its own partials); after a
// second barrier it pairs it with the partner half and adds the rounded coefficients it keeps into
// the accumulator (split: LDS integer atomics, the hi half shifted by the split, exact mod 2^32 in
// either order; FP64: each coefficient has one writer).  A step's critical path is two forward
// half-transforms and one inverse half-transform instead of ten.  FP64 outputs equal k_boot2w's
// bit for bit: both are the exact integers the final rounding recovers.  For launches of at most
// one bootstrap per CU (modes 0, 1 and 3).
// ---------------------------------------------------------------------------------------------
template <int L, int BG, int P>
__global__ void __launch_bounds__(64 * (P ? 8 : 2 * L)) k_boot2wlat(BootKArgs A) {
    constexpr int N = W2_N, M = W2_M, R1 = 8, NC = 32, KPL = 2 * L;
    constexpr int K = P ? 4 : 2;          // outputs per parity
    constexpr int NW = P ? 8 : 2 * L;      // wavefronts: 2l forward, 2K inverse
    static_assert(2 * K <= NW && 2 * L <= NW, "every forward and inverse job has a wavefront");
    constexpr int XB = xbuf_elems<W2_H>();
    constexpr int T2S = 9;
    constexpr uint32_t OFF = decomp_offset<L, BG>();
    using T = typename std::conditional<P == 1, float, double>::type;
    // accumulator with mirrored ends (w2_rotate)
    constexpr int G = W2_G_SPLIT, MW = w2_mirror<G>();
    __shared__ int32_t s_acc_m[2 * N + 2 * MW];
    int32_t* const s_acc = s_acc_m + MW;
    __shared__ int16_t s_bara[kMaxLweN];
    __shared__ double s_x[NW][XB];              // transform exchanges: C2<float>[XB] or double[XB] (half width)
    __shared__ C2<T> s_part[2][K][L][W2_H];     // [parity][output][level][frequency]; then the output exchange
    __shared__ C2<double> s_t2[P ? 1 : 8 * T2S];
    static_assert(sizeof(s_part[0][0]) >= W2_H * sizeof(C2<double>), "an output half fits over its partials");
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int v = w & 1;
    const int g = blockIdx.x;
    if (w == 0) {
        boot_prologue<N>(A, g, lane, s_acc, s_bara);
        w2_mirror_fill<G>(s_acc, lane);   // lane-local copies, no fence (see k_boot2w)
    }
    if constexpr (P == 0) {
        if (threadIdx.x < 64) s_t2[(threadIdx.x >> 3) * T2S + (threadIdx.x & 7)] = ((const C2<double>*)A.t2)[threadIdx.x];
    }
    __syncthreads();
    constexpr size_t ROWSET = (size_t)KPL * 2 * M;
    using TW = typename std::conditional<P == 1, LaneTw<float, 8>, LaneTwS<double, 8>>::type;
    TW tw;
    [[maybe_unused]] f16x8_t a2[2];
    if constexpr (P == 1) {
#if TFHE_W2LAT_FWD_MFMA
        // the pass-1 twiddles in pass 2's layout (forward and inverse) and this lane's entries of the
        // two FP16 pieces of parity v's pass-1 matrix
        load_lane_tw(tw, (const C2<float>*)A.fwd_t2 + v * 64 * 8, (const C2<float>*)A.t2, lane);
        a2[0] = ((const f16x8_t*)A.fwd_a16)[(v * 64 + lane) * 2];
        a2[1] = ((const f16x8_t*)A.fwd_a16)[(v * 64 + lane) * 2 + 1];
#else
        load_lane_tw(tw, (const C2<float>*)A.t1 + v * 64 * 8, (const C2<float>*)A.t2, lane);
#endif
    } else {
#pragma unroll
        for (int r = 0; r < 8; ++r) tw.t1[r] = ((const C2<double>*)A.t1)[(v * 64 + lane) * 8 + r];
        tw.t2 = s_t2 + (lane & 7) * T2S;
    }
    const C2<double>* rw64 = (const C2<double>*)A.t2w_rw64 + v * 8;     // FP64: forward and untwist
    const C2<float>* rw32 = (const C2<float>*)A.t2w_rw32 + v * 8;       // split: forward
    const C2<double>* rws = (const C2<double>*)A.untw64 + v * 8;        // split: untwist, 1/1024 folded in
    const float hc = v ? -0.70710678118654752440f : 0.70710678118654752440f;
    const int n = A.n;
    float sink = 0.f, warm[4] = {0.f, 0.f, 0.f, 0.f};
    for (int i = 0; i < n; ++i) {
        const int a = __builtin_amdgcn_readfirstlane((int)s_bara[i]);
        if (a == 0) continue;   // tfhe_blindRotate_FFT.cpp:114; uniform over the workgroup
        if (w < 2 * L) {
            const int lvl = w >> 1;
            C2<T> acc[K][R1];
#pragma unroll
            for (int k = 0; k < K; ++k)
#pragma unroll
                for (int r = 0; r < R1; ++r) acc[k][r] = C2<T>{(T)0, (T)0};
            int t2o = 0;
            asm volatile("" : "+v"(t2o));
            TW twl = tw;
            if constexpr (P == 0) twl.t2 = tw.t2 + t2o;
            if constexpr (P == 1) {
                // the two rows (0, lvl) and (1, lvl), software-pipelined: each row's first key
                // polynomial is requested before its transform, the second after it, and polynomial
                // 1 is rotated and its digits prepared while row (0, lvl)'s second polynomial is in
                // flight
                uint32_t u[NC];
#if TFHE_W2LAT_FWD_MFMA
                // x[aa] carries the FP16 digit pairs (P0_aa, P1_aa) until its transform
                (void)hc;
                (void)rw32;
                auto rotate = [&](int q) { w2_rotate<G, true, OFF, false>(u, s_acc, q, a, lane); };
                auto prep = [&](C2<float>* x) {
#pragma unroll
                    for (int aa = 0; aa < R1; ++aa)
                        x[aa] = C2<float>{__uint_as_float(digits8_f16x2(u[aa], u[aa + 16], lvl)),
                                          __uint_as_float(digits8_f16x2(u[aa + 8], u[aa + 24], lvl))};
                };
                auto fwd = [&](C2<float>* x) {
                    uint32_t p0[R1], p1[R1];
#pragma unroll
                    for (int aa = 0; aa < R1; ++aa) {
                        p0[aa] = __float_as_uint(x[aa].x);
                        p1[aa] = __float_as_uint(x[aa].y);
                    }
                    fft512_fwd2w_mfma<false>(x, p0, p1, a2, twl.t1, twl.t2, (C2<float>*)s_x[w], lane);
                };
#else
                auto rotate = [&](int q) { w2_rotate<G, true, OFF>(u, s_acc, q, a, lane); };
                auto prep = [&](C2<float>* x) {
#pragma unroll
                    for (int aa = 0; aa < R1; ++aa)
                        x[aa] = w2_prep32((float)digit_sx<BG>(u[aa], lvl), (float)digit_sx<BG>(u[aa + 16], lvl),
                                          (float)digit_sx<BG>(u[aa + 8], lvl), (float)digit_sx<BG>(u[aa + 24], lvl),
                                          hc, rw32[aa], aa == 0);
                };
                auto fwd = [&](C2<float>* x) { fft_fwd<R1>(x, twl, s_x[w], lane); };
#endif
                C2<float> x[R1];
                rotate(0);
                prep(x);
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    const float4* kp = (const float4*)A.bk + (size_t)i * ROWSET + (size_t)((q * L + lvl) * 2) * M +
                                       v * W2_H + lane;
                    float4 k0[R1];
#pragma unroll
                    for (int r = 0; r < R1; ++r) k0[r] = kp[r * 64];
                    fwd(x);
#pragma unroll
                    for (int r = 0; r < R1; ++r) {
                        cmac(acc[0][r], x[r], C2<float>{k0[r].x, k0[r].y});
                        cmac(acc[1][r], x[r], C2<float>{k0[r].z, k0[r].w});
                    }
                    float4 k1[R1];
#pragma unroll
                    for (int r = 0; r < R1; ++r) k1[r] = kp[M + r * 64];
                    C2<float> xn[R1];
                    if (q == 0) {
                        rotate(1);
                        prep(xn);
                    }
#pragma unroll
                    for (int r = 0; r < R1; ++r) {
                        cmac(acc[2][r], x[r], C2<float>{k1[r].x, k1[r].y});
                        cmac(acc[3][r], x[r], C2<float>{k1[r].z, k1[r].w});
                    }
                    if (q == 0) {
#pragma unroll
                        for (int r = 0; r < R1; ++r) x[r] = xn[r];
                    }
                }
            } else {
                // the two rows pipelined as on the split path: the first key polynomial before the
                // transform, then EK1 of the second polynomial's values before polynomial 1's
                // rotation and digits (all eight fit too, but ran slower), the rest after them
                constexpr int EK1 = 4;
                uint32_t u[NC];
                auto rotate = [&](int q) { w2_rotate<G, true, OFF>(u, s_acc, q, a, lane); };
                auto prep = [&](C2<double>* x) {
#pragma unroll
                    for (int aa = 0; aa < R1; ++aa) {
                        const C2<double> p0{(double)digit_sx<BG>(u[aa], lvl), (double)digit_sx<BG>(u[aa + 16], lvl)};
                        const C2<double> p1{(double)digit_sx<BG>(u[aa + 8], lvl), (double)digit_sx<BG>(u[aa + 24], lvl)};
                        x[aa] = w2_prep(p0, p1, rw64[aa], v, aa == 0);
                    }
                };
                C2<double> x[R1];
                rotate(0);
                prep(x);
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    const C2<double>* kp = (const C2<double>*)A.bk + (size_t)i * ROWSET +
                                           (size_t)((q * L + lvl) * 2) * M + v * W2_H + lane;
                    C2<double> k0[R1];
#pragma unroll
                    for (int r = 0; r < R1; ++r) k0[r] = kp[r * 64];
                    fft_fwd<R1, double, true, false>(x, twl, s_x[w], lane);
#pragma unroll
                    for (int r = 0; r < R1; ++r) cmac(acc[0][r], x[r], k0[r]);
                    C2<double> k1[EK1];
#pragma unroll
                    for (int r = 0; r < EK1; ++r) k1[r] = kp[M + r * 64];
                    C2<double> xn[R1];
                    if (q == 0) {
                        rotate(1);
                        prep(xn);
                    }
#pragma unroll
                    for (int r = 0; r < R1; ++r) cmac(acc[1][r], x[r], r < EK1 ? k1[r < EK1 ? r : 0] : kp[M + r * 64]);
                    if (q == 0) {
#pragma unroll
                        for (int r = 0; r < R1; ++r) x[r] = xn[r];
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < K; ++k)
#pragma unroll
                for (int r = 0; r < R1; ++r) s_part[v][k][lvl][r * 64 + lane] = acc[k][r];
            // a single bootstrap shares its key rows with nobody, so every row would come from
            // memory: touch every 128-byte line of this wavefront's two rows of the next step now,
            // a step before they are read.  The loaded words are only folded into `sink` at the
            // next step, so nothing waits for them here.
            if (i + 1 < n) {
                const float4* nk = (const float4*)A.bk + (size_t)(i + 1) * ROWSET;
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    const float* rp = (const float*)(nk + (size_t)((q * L + lvl) * 2) * M + v * W2_H);
                    sink += warm[2 * q] + warm[2 * q + 1];
                    warm[2 * q] = rp[lane * 32];
                    warm[2 * q + 1] = rp[W2_M * 4 + lane * 32];
                }
            }
        }
        __syncthreads();
        const int k = w >> 1;
        C2<double> own[R1];
        if (w < 2 * K) {
            int t2o = 0;
            asm volatile("" : "+v"(t2o));
            TW twl = tw;
            if constexpr (P == 0) twl.t2 = tw.t2 + t2o;
            C2<T> o[R1];
#pragma unroll
            for (int r = 0; r < R1; ++r) {
                o[r] = s_part[v][k][0][r * 64 + lane];
#pragma unroll
                for (int t = 1; t < L; ++t) o[r] = o[r] + s_part[v][k][t][r * 64 + lane];
            }
            C2<double>* xo = (C2<double>*)&s_part[v][k][0][0];
            if constexpr (P == 1) {
#if TFHE_W2LAT_FWD_MFMA
                fft512_inv2w_t1p(o, twl.t1, twl.t2, (C2<float>*)s_x[w], lane);
#else
                fft_inv<R1>(o, twl, s_x[w], lane);
#endif
#pragma unroll
                for (int aa = 0; aa < R1; ++aa) {
                    own[aa] = cmulc(C2<double>{(double)o[aa].x, (double)o[aa].y}, rws[aa]);
                    xo[aa * 64 + lane] = own[aa];
                }
            } else {
                fft_inv<R1, double, true, false>(o, twl, s_x[w], lane);
#pragma unroll
                for (int aa = 0; aa < R1; ++aa) {
                    own[aa] = aa ? cmulc(o[aa], rw64[aa]) : o[aa];
                    xo[aa * 64 + lane] = own[aa];
                }
            }
        }
        __syncthreads();
        if (w < 2 * K) {
            const C2<double>* xp = (const C2<double>*)&s_part[1 - v][k][0][0];
            if constexpr (P == 1) {
                const int j = k >> 1, sh = (k & 1) ? 0 : A.split_bits;   // k = 2j + (0 hi, 1 lo)
                int32_t* accp = s_acc + j * N;
#pragma unroll
                for (int aa = 0; aa < R1; ++aa) {
                    const C2<double> z = w2_pair_own(own[aa], xp[aa * 64 + lane], v);
                    const int32_t d0 = (int32_t)(round_mod32(z.x) << sh), d1 = (int32_t)(round_mod32(z.y) << sh);
                    const int c0 = w2_coef(v, lane, aa, 0), c1 = w2_coef(v, lane, aa, 1);
                    atomicAdd(accp + c0, d0);
                    atomicAdd(accp + c1, d1);
                    // the mirrors take the same increments (w2_mirror_update's coefficient sets)
                    if (j == 0 && v == 1 && aa >= 8 - MW / 64) atomicAdd(accp + c1 - N, d1);
                    if (j == 1 && v == 0 && aa < MW / 64) atomicAdd(accp + c0 + N, d0);
                }
            } else {
                int32_t* accp = s_acc + k * N;
#pragma unroll
                for (int aa = 0; aa < R1; ++aa) {
                    const C2<double> y = xp[aa * 64 + lane];
                    const C2<double> z = w2_pair(own[aa], y, v);
                    const int j0 = w2_coef(v, lane, aa, 0), j1 = w2_coef(v, lane, aa, 1);
                    const int32_t n0 = (int32_t)((uint32_t)accp[j0] + round_mod32(z.x));
                    const int32_t n1 = (int32_t)((uint32_t)accp[j1] + round_mod32(z.y));
                    accp[j0] = n0;
                    accp[j1] = n1;
                    w2_mirror_update<G>(s_acc, k, v, aa, k * N + j0, k * N + j1, n0, n1);
                }
            }
        }
        __syncthreads();
    }
    asm volatile("" ::"v"(sink + warm[0] + warm[1] + warm[2] + warm[3]));
    if (w == 0) boot_epilogue<N>(A, g, lane, s_acc);
}

// Key preparation at N = 2048: one two-wavefront workgroup per key polynomial, the same prep and
// 512-point transform as the bootstrap (FP64), stored as [parity][register][lane].
__global__ void __launch_bounds__(128) k_keyprep2w(const int32_t* __restrict__ bk, C2<double>* __restrict__ out64,
                                                   float4* __restrict__ out32, int split_bits,
                                                   const C2<double>* __restrict__ t1w, const C2<double>* __restrict__ t2,
                                                   const C2<double>* __restrict__ rw) {
    constexpr int N = W2_N, M = W2_M, R1 = 8;
    __shared__ C2<double> s_buf[2][xbuf_elems<W2_H>()];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const size_t poly = blockIdx.x;
    const int32_t* p = bk + poly * N;
    LaneTw<double, R1> tw;
    load_lane_tw(tw, t1w + wave * 64 * 8, t2, lane);
    const C2<double>* rwv = rw + wave * 8;
    const size_t obase = poly * M + wave * W2_H;
    C2<double> v[R1];
    if (out64) {
#pragma unroll
        for (int aa = 0; aa < R1; ++aa) {
            const int j = lane + 64 * aa;
            v[aa] = w2_prep(C2<double>{(double)p[j], (double)p[j + 1024]}, C2<double>{(double)p[j + 512], (double)p[j + 1536]},
                            rwv[aa], wave, aa == 0);
        }
        fft_fwd<R1>(v, tw, s_buf[wave], lane);
#pragma unroll
        for (int r = 0; r < R1; ++r) out64[obase + r * 64 + lane] = v[r];
    }
    if (out32) {
        C2<double> h[R1];
        const int64_t half = (int64_t)1 << (split_bits - 1);
        auto hi = [&](int64_t x) { return (x + half) >> split_bits; };
#pragma unroll
        for (int aa = 0; aa < R1; ++aa) {
            const int j = lane + 64 * aa;
            const int64_t x0 = p[j], x1 = p[j + 1024], x2 = p[j + 512], x3 = p[j + 1536];
            const int64_t h0 = hi(x0), h1 = hi(x1), h2 = hi(x2), h3 = hi(x3);
            h[aa] = w2_prep(C2<double>{(double)h0, (double)h1}, C2<double>{(double)h2, (double)h3}, rwv[aa], wave, aa == 0);
            v[aa] = w2_prep(C2<double>{(double)(x0 - (h0 << split_bits)), (double)(x1 - (h1 << split_bits))},
                            C2<double>{(double)(x2 - (h2 << split_bits)), (double)(x3 - (h3 << split_bits))}, rwv[aa],
                            wave, aa == 0);
        }
        fft_fwd<R1>(h, tw, s_buf[wave], lane);
        fft_fwd<R1>(v, tw, s_buf[wave], lane);
#pragma unroll
        for (int r = 0; r < R1; ++r)
            out32[obase + r * 64 + lane] = make_float4((float)h[r].x, (float)h[r].y, (float)v[r].x, (float)v[r].y);
    }
}

// Transform test kernel at N = 2048 (two wavefronts): forward int32[N] -> F_k natural order, or
// inverse F natural -> polynomial (FP64 values), with the bootstrap's prep and output pairing of
// each path (P = 1: split, whose wavefronts untwist their own outputs with rws64 and exchange them).
// inverse: 0 forward, 1 inverse; for P = 1 (split) 2 is the forward of digit inputs (|p_j| < 2048)
// with the matrix-core form (fft512_fwd2w_mfma), and the inverse is the one the bootstrap kernel
// runs (fft512_inv2w_t1p with the pass-2-layout table t1p when TFHE_W2_FWD_MFMA is 1).
template <int P>
__global__ void __launch_bounds__(128) k_fft_test2w(double* __restrict__ out, const void* __restrict__ in, int inverse,
                                                    const void* __restrict__ t1w, const void* __restrict__ t2,
                                                    const C2<float>* __restrict__ rw32,
                                                    const C2<double>* __restrict__ rw64,
                                                    const C2<double>* __restrict__ rws64,
                                                    const void* __restrict__ w2_a16, const void* __restrict__ t1p) {
    using T = typename std::conditional<P == 1, float, double>::type;
    constexpr int N = W2_N, R1 = 8;
    __shared__ C2<double> s_buf[2][xbuf_elems<W2_H>()];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const size_t g = blockIdx.x;
    LaneTw<T, R1> tw;
    constexpr bool MF = P == 1 && TFHE_W2_FWD_MFMA;
    const bool use_t1p = P == 1 && (inverse == 2 || (inverse == 1 && MF));
    load_lane_tw(tw, (const C2<T>*)(use_t1p ? t1p : t1w) + wave * 64 * 8, (const C2<T>*)t2, lane);
    C2<T> v[R1];
    double* o = out + g * N;
    if (inverse == 2) {
        if constexpr (P == 1) {
            const int32_t* p = (const int32_t*)in + g * N;
            uint32_t x0[R1], x1[R1];
#pragma unroll
            for (int aa = 0; aa < R1; ++aa) {
                // the coefficients of register aa as the bootstrap kernel reads them
                const int j = TFHE_W2_BLAYOUT ? bl_lane(lane) + bl_off(aa) : lane + 64 * aa;
                x0[aa] = pack_f16x2((float)p[j], (float)p[j + 1024]);
                x1[aa] = pack_f16x2((float)p[j + 512], (float)p[j + 1536]);
            }
            f16x8_t a2[2];
            a2[0] = ((const f16x8_t*)w2_a16)[(wave * 64 + lane) * 2];
            a2[1] = ((const f16x8_t*)w2_a16)[(wave * 64 + lane) * 2 + 1];
            fft512_fwd2w_mfma<TFHE_W2_BLAYOUT>(v, x0, x1, a2, tw.t1, tw.t2, (C2<float>*)s_buf[wave], lane);
#pragma unroll
            for (int r = 0; r < R1; ++r) {
                const int k = 2 * (lane + 64 * r) + wave;
                o[2 * k] = v[r].x;
                o[2 * k + 1] = v[r].y;
            }
        }
    } else if (!inverse) {
        const int32_t* p = (const int32_t*)in + g * N;
#pragma unroll
        for (int aa = 0; aa < R1; ++aa) {
            const int j = lane + 64 * aa;
            if constexpr (P == 1)
                v[aa] = w2_prep32((float)p[j], (float)p[j + 1024], (float)p[j + 512], (float)p[j + 1536],
                                  wave ? -0.70710678118654752440f : 0.70710678118654752440f, rw32[wave * 8 + aa],
                                  aa == 0);
            else
                v[aa] = w2_prep(C2<T>{(T)p[j], (T)p[j + 1024]}, C2<T>{(T)p[j + 512], (T)p[j + 1536]},
                                rw64[wave * 8 + aa], wave, aa == 0);
        }
        fft_fwd<R1>(v, tw, s_buf[wave], lane);
#pragma unroll
        for (int r = 0; r < R1; ++r) {
            const int k = 2 * (lane + 64 * r) + wave;
            o[2 * k] = v[r].x;
            o[2 * k + 1] = v[r].y;
        }
    } else {
        const double* F = (const double*)in + g * N;
#pragma unroll
        for (int r = 0; r < R1; ++r) {
            const int k = 2 * (lane + 64 * r) + wave;
            v[r] = C2<T>{(T)F[2 * k], (T)F[2 * k + 1]};
        }
        if constexpr (MF) fft512_inv2w_t1p((C2<float>*)v, tw.t1, tw.t2, (C2<float>*)s_buf[wave], lane);
        else fft_inv<R1>(v, tw, s_buf[wave], lane);
        __syncthreads();
        C2<double>* xo = (C2<double>*)s_buf[wave];
        const C2<double>* xp = (const C2<double>*)s_buf[1 - wave];
        if constexpr (P == 1) {
            // as the split bootstrap step: own half untwisted in FP64, exchanged rounded to FP32
            C2<double> own[R1];
            C2<float>* xo32 = (C2<float>*)xo;
            const C2<float>* xp32 = (const C2<float>*)xp;
#pragma unroll
            for (int aa = 0; aa < R1; ++aa) {
                own[aa] = cmulc(C2<double>{(double)v[aa].x, (double)v[aa].y}, rws64[wave * 8 + aa]);
                xo32[aa * 64 + lane] = C2<float>{(float)own[aa].x, (float)own[aa].y};
            }
            __syncthreads();
#pragma unroll
            for (int aa = 0; aa < R1; ++aa) {
                const C2<float> y = xp32[aa * 64 + lane];
                const C2<double> z = w2_pair_own(own[aa], C2<double>{(double)y.x, (double)y.y}, wave);
                o[w2_coef(wave, lane, aa, 0)] = z.x;
                o[w2_coef(wave, lane, aa, 1)] = z.y;
            }
        } else {
#pragma unroll
            for (int aa = 0; aa < R1; ++aa) xo[aa * 64 + lane] = C2<double>{(double)v[aa].x, (double)v[aa].y};
            __syncthreads();
#pragma unroll
            for (int aa = 0; aa < R1; ++aa) {
                const C2<double> own = cmulc(C2<double>{(double)v[aa].x, (double)v[aa].y}, rw64[wave * 8 + aa]);
                const C2<double> oth = cmulc(xp[aa * 64 + lane], rw64[(1 - wave) * 8 + aa]);
                const C2<double> z = w2_pair(own, oth, wave);
                o[w2_coef(wave, lane, aa, 0)] = z.x;
                o[w2_coef(wave, lane, aa, 1)] = z.y;
            }
        }
    }
}

// =============================================================================================
// Split-FP32 bootstrap with the pointwise products on the matrix cores (launch form 6, N = 1024).
//
// The per-frequency products of the external product, out[o][f] += dec[f] * K[o][f], use the same
// key value for every bootstrap, so they form a contraction (and can run on the matrix cores) only
// across bootstraps.  A wavefront here runs 4 bootstraps: lane 4 l + b works on bootstrap b
// (l < 16), and each bootstrap's 512-point transform is spread over its 16 lanes, 32 points each:
//   forward  lane l, register r holds z_j, j = 16 r + l, taken in as (d_j + i d_{j+512}) w^{16 r}
//            (w = e^{i pi / 1024}); pass 1: a 32-point DFT over r gives Y_l[k1]; pass-2 twiddle
//            T[l][k1] = e^{i pi l (4 k1 + 1) / 1024}, which also carries the lane part w^l of the
//            twist; one LDS exchange gives lane l' the values k1 = l' + 16 s of every l; pass 2: a
//            16-point DFT over l for each s.  Register 16 s + k2 of lane l' then holds F_f,
//            f = l' + 16 s + 32 k2, in the transform convention of the other kernels.
//   products for every register, v_mfma_f32_4x4x1_16b_f32 with block l' (16 frequencies), rows =
//            the four components (hi.re, hi.im, lo.re, lo.im) of one output polynomial's key value
//            and columns = the four bootstraps: D[c][4 l' + b] += A[4 l' + c] * B[4 l' + b], with B
//            the real or imaginary part of the digit spectrum (the register itself) and A the key
//            arranged for it: K's components for the real part, (-K.im, K.re) per half for the
//            imaginary part.  The sums land in the transform's own layout, lane 4 l' + b.
//   inverse  the passes in reverse with conjugate twiddles; the uniform untwist conj(w^{16 r}) /
//            512 and the rounding run in FP64, as in the other split kernels.
// Key (k_keyprep_q4): per LWE index, row and register one float4 per lane, {A_re(K_0), A_im(K_0),
// A_re(K_1), A_im(K_1)} (component lane & 3 of each), 1 KB per register, 192 KB per row set; one
// load serves four bootstraps.
// LDS: the 4 accumulators (32 KB) and a 4.25 KB exchange buffer: four wavefronts per CU, one per
// SIMD.  The modulus-switched rotations are recomputed from the inputs at each step.
// =============================================================================================
// e^{i pi r / 64}, r < 32: the uniform part w^{16 r} of the forward twist (w = e^{i pi / 1024})
constexpr float kQ4TwRe[32] = {
    1.000000000e+00f, 9.987954562e-01f, 9.951847267e-01f, 9.891765100e-01f, 9.807852804e-01f, 9.700312532e-01f,
    9.569403357e-01f, 9.415440652e-01f, 9.238795325e-01f, 9.039892931e-01f, 8.819212643e-01f, 8.577286100e-01f,
    8.314696123e-01f, 8.032075315e-01f, 7.730104534e-01f, 7.409511254e-01f, 7.071067812e-01f, 6.715589548e-01f,
    6.343932842e-01f, 5.956993045e-01f, 5.555702330e-01f, 5.141027442e-01f, 4.713967368e-01f, 4.275550934e-01f,
    3.826834324e-01f, 3.368898534e-01f, 2.902846773e-01f, 2.429801799e-01f, 1.950903220e-01f, 1.467304745e-01f,
    9.801714033e-02f, 4.906767433e-02f
};
constexpr float kQ4TwIm[32] = {
    0.000000000e+00f, 4.906767433e-02f, 9.801714033e-02f, 1.467304745e-01f, 1.950903220e-01f, 2.429801799e-01f,
    2.902846773e-01f, 3.368898534e-01f, 3.826834324e-01f, 4.275550934e-01f, 4.713967368e-01f, 5.141027442e-01f,
    5.555702330e-01f, 5.956993045e-01f, 6.343932842e-01f, 6.715589548e-01f, 7.071067812e-01f, 7.409511254e-01f,
    7.730104534e-01f, 8.032075315e-01f, 8.314696123e-01f, 8.577286100e-01f, 8.819212643e-01f, 9.039892931e-01f,
    9.238795325e-01f, 9.415440652e-01f, 9.569403357e-01f, 9.700312532e-01f, 9.807852804e-01f, 9.891765100e-01f,
    9.951847267e-01f, 9.987954562e-01f
};
// the same in FP64 with the inverse transform's 1/512 folded in (the untwist multiplies by the conjugate)
constexpr double kQ4UntwRe[32] = {
    1.95312500000000000e-03, 1.95077237540072735e-03, 1.94372016928163463e-03, 1.93198537102496292e-03,
    1.91559625078755943e-03, 1.89459229139559370e-03, 1.86902409322697036e-03, 1.83895325231058751e-03,
    1.80445221193610691e-03, 1.76560408813172527e-03, 1.72250246943038096e-03, 1.67525119140678148e-03,
    1.62396408652840866e-03, 1.56876470992313465e-03, 1.50978604172409569e-03, 1.44717016670890450e-03,
    1.38106793200497573e-03, 1.31163858368558268e-03, 1.23904938313212009e-03, 1.16347520408678412e-03,
    1.08509811136641072e-03, 1.00410692225238606e-03, 9.20696751613276970e-04, 8.35068541856019914e-04,
    7.47428578838066088e-04, 6.57987994906679787e-04, 5.66962260262621740e-04, 4.74570663873562465e-04,
    3.81035785187750647e-04, 2.86582957920628414e-04, 1.91439727206173379e-04, 9.58353014207385273e-05
};
constexpr double kQ4UntwIm[32] = {
    0.00000000000000000e+00, 9.58353014207383104e-05, 1.91439727206173054e-04, 2.86582957920628414e-04,
    3.81035785187750485e-04, 4.74570663873562248e-04, 5.66962260262621740e-04, 6.57987994906679787e-04,
    7.47428578838065980e-04, 8.35068541856019697e-04, 9.20696751613276645e-04, 1.00410692225238606e-03,
    1.08509811136641050e-03, 1.16347520408678390e-03, 1.23904938313212009e-03, 1.31163858368558268e-03,
    1.38106793200497551e-03, 1.44717016670890450e-03, 1.50978604172409569e-03, 1.56876470992313444e-03,
    1.62396408652840866e-03, 1.67525119140678148e-03, 1.72250246943038074e-03, 1.76560408813172527e-03,
    1.80445221193610691e-03, 1.83895325231058751e-03, 1.86902409322697058e-03, 1.89459229139559370e-03,
    1.91559625078755943e-03, 1.93198537102496292e-03, 1.94372016928163441e-03, 1.95077237540072735e-03
};
// W_32^m = e^{2 pi i m / 32}, m < 32
constexpr float kW32Re[32] = {
    1.000000000e+00f, 9.807852804e-01f, 9.238795325e-01f, 8.314696123e-01f, 7.071067812e-01f, 5.555702330e-01f,
    3.826834324e-01f, 1.950903220e-01f, 0.000000000e+00f, -1.950903220e-01f, -3.826834324e-01f, -5.555702330e-01f,
    -7.071067812e-01f, -8.314696123e-01f, -9.238795325e-01f, -9.807852804e-01f, -1.000000000e+00f,
    -9.807852804e-01f, -9.238795325e-01f, -8.314696123e-01f, -7.071067812e-01f, -5.555702330e-01f,
    -3.826834324e-01f, -1.950903220e-01f, 0.000000000e+00f, 1.950903220e-01f, 3.826834324e-01f, 5.555702330e-01f,
    7.071067812e-01f, 8.314696123e-01f, 9.238795325e-01f, 9.807852804e-01f
};
constexpr float kW32Im[32] = {
    0.000000000e+00f, 1.950903220e-01f, 3.826834324e-01f, 5.555702330e-01f, 7.071067812e-01f, 8.314696123e-01f,
    9.238795325e-01f, 9.807852804e-01f, 1.000000000e+00f, 9.807852804e-01f, 9.238795325e-01f, 8.314696123e-01f,
    7.071067812e-01f, 5.555702330e-01f, 3.826834324e-01f, 1.950903220e-01f, 0.000000000e+00f, -1.950903220e-01f,
    -3.826834324e-01f, -5.555702330e-01f, -7.071067812e-01f, -8.314696123e-01f, -9.238795325e-01f,
    -9.807852804e-01f, -1.000000000e+00f, -9.807852804e-01f, -9.238795325e-01f, -8.314696123e-01f,
    -7.071067812e-01f, -5.555702330e-01f, -3.826834324e-01f, -1.950903220e-01f
};
typedef float q4v4f __attribute__((ext_vector_type(4)));
constexpr int Q4_NB = 4;                       // bootstraps per wavefront
constexpr uint32_t Q4_ROWSET_B = 6u * 32u * 64u * 16u;   // bytes of one LWE index's row set (l = 3)

template <int S>
__device__ __forceinline__ void q4_dft4(C2<float>& x0, C2<float>& x1, C2<float>& x2, C2<float>& x3) {
    const C2<float> a0 = x0 + x2, a1 = x0 - x2, b0 = x1 + x3, b1 = mul_i<S>(x1 - x3);
    x0 = a0 + b0;
    x2 = a0 - b0;
    x1 = a1 + b1;
    x3 = a1 - b1;
}

// X_k at x[q4_slot(k)] after q4_dft32
__host__ __device__ constexpr int q4_slot(int k) { return 8 * (k & 3) + (k >> 2); }

// 32-point DFT X_k = sum_n x_n e^{S 2 pi i n k / 32}, natural order in: n = 8 n1 + n2, k = k1 + 4 k2;
// 4-point DFTs over n1, twiddles W_32^{S n2 k1} fused into the 8-point DFTs over n2
template <int S>
__device__ __forceinline__ void q4_dft32(C2<float>* x) {
#pragma unroll
    for (int n2 = 0; n2 < 8; ++n2) q4_dft4<S>(x[n2], x[8 + n2], x[16 + n2], x[24 + n2]);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        C2<float> w[8];
#pragma unroll
        for (int n2 = 0; n2 < 8; ++n2) w[n2] = C2<float>{kW32Re[(n2 * k1) & 31], kW32Im[(n2 * k1) & 31]};
        dft8_tw<S, (S < 0), true>(x + 8 * k1, w);
    }
}

// The exchange moves the values of one s-half (k1 = c + 16 s, c < 16) at a time, real parts then
// imaginary parts: element (c, source lane x = 4 l + b) at 68 c + x.  Writes (one c, every lane)
// and reads (one source l of c = l') meet 64 distinct banks, and both addresses are a per-lane
// base plus an immediate offset: 1088 floats (4352 B) of buffer.
constexpr int Q4_XROW = 68, Q4_XBUF = 16 * Q4_XROW;

// v[q4_slot(k1)] = Y[k1] (k1 < 32) -> v[q4_slot(16 s + l)] = Y_l[l' + 16 s] (this lane is l'): each
// half s returns in the registers it left from, so no value of the other half is overwritten, and
// within a half the real parts travel before the imaginary ones.
__device__ __forceinline__ void q4_exchange_fwd(C2<float>* v, float* sx, int lane) {
    const int ml = lane >> 2, mb = lane & 3;
    float* wb = sx + lane;
    const float* rb = sx + Q4_XROW * ml + mb;
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int part = 0; part < 2; ++part) {
#pragma unroll
            for (int c = 0; c < 16; ++c) wb[Q4_XROW * c] = part ? v[q4_slot(c + 16 * s)].y : v[q4_slot(c + 16 * s)].x;
            wave_sync();
#pragma unroll
            for (int src = 0; src < 16; ++src) {
                const float val = rb[4 * src];
                if (part) v[q4_slot(src + 16 * s)].y = val;
                else v[q4_slot(src + 16 * s)].x = val;
            }
            wave_sync();
        }
}

// the reverse: v[16 s + l] (values of k1 = l' + 16 s from every source l) -> v[k1], natural order
// (half s returns in the registers it left from)
__device__ __forceinline__ void q4_exchange_inv(C2<float>* v, float* sx, int lane) {
    const int ml = lane >> 2, mb = lane & 3;
    float* wb = sx + Q4_XROW * ml + mb;
    const float* rb = sx + lane;
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int part = 0; part < 2; ++part) {
#pragma unroll
            for (int src = 0; src < 16; ++src) wb[4 * src] = part ? v[16 * s + src].y : v[16 * s + src].x;
            wave_sync();
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                const float val = rb[Q4_XROW * c];
                if (part) v[16 * s + c].y = val;
                else v[16 * s + c].x = val;
            }
            wave_sync();
        }
}

// Pass-2 twiddles T[i][l' + 16 s] = e^{i pi i (4 (l' + 16 s) + 1) / 1024} (i < 16 the source lane),
// table [l'][s][i] of float2 (4 KB, DevTwiddles::q4tw), read at each use through the cache rather
// than held in 64 registers.  tt: this lane's row, tt[16 s + i].
__device__ __forceinline__ C2<float> q4_tw(const C2<float>* __restrict__ tt, int k) {
    return tt[k];
}

// an opaque copy of the table row pointer per transform: the loads of a transform cannot be hoisted
// out of the blind-rotation loop (which would hold 60 registers across it)
__device__ __forceinline__ const C2<float>* q4_opaque(const C2<float>* p) {
    asm volatile("" : "+v"(p));
    return p;
}

// Forward transform, 4 bootstraps per wavefront.  In: v[r] = (d_{16 r + l}, d_{16 r + l + 512}).
// Out: v[16 s + k2] = F_{l + 16 s + 32 k2}.
__device__ __forceinline__ void q4_fwd(C2<float>* v, const C2<float>* __restrict__ tt, float* sx, int lane) {
#pragma unroll
    for (int r = 1; r < 32; ++r) v[r] = cmul(v[r], C2<float>{kQ4TwRe[r], kQ4TwIm[r]});
    q4_dft32<+1>(v);
    q4_exchange_fwd(v, sx, lane);
    C2<float> h[2][16];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const C2<float>* t = q4_opaque(tt);
#pragma unroll
        for (int i = 0; i < 16; ++i) h[s][i] = i ? cmul(v[q4_slot(16 * s + i)], q4_tw(t, 16 * s + i)) : v[q4_slot(16 * s)];
        dft16<+1>(h[s]);
        __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int k2 = 0; k2 < 16; ++k2) v[16 * s + k2] = h[s][k2];
}

// Inverse of q4_fwd without the 1/512 and the uniform untwist: in v[16 s + k2] = F_{l + 16 s + 32 k2};
// out v[q4_slot(r)] = 512 z_{16 r + l} w^{-16 r} (z_j = p_j + i p_{j+512}).
__device__ __forceinline__ void q4_inv(C2<float>* v, const C2<float>* __restrict__ tt, float* sx, int lane) {
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const C2<float>* t = q4_opaque(tt);
        dft16<-1>(v + 16 * s);
#pragma unroll
        for (int i = 1; i < 16; ++i) v[16 * s + i] = cmulc(v[16 * s + i], q4_tw(t, 16 * s + i));
        __builtin_amdgcn_sched_barrier(0);
    }
    q4_exchange_inv(v, sx, lane);
    q4_dft32<-1>(v);
}

template <int L, int BG>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1))) k_boot32q(BootKArgs A) {
    constexpr int N = 1024, KPL = 2 * L;
    static_assert(KPL == 6, "the key row set layout assumes l = 3");
    constexpr uint32_t OFF = decomp_offset<L, BG>();
    constexpr int SH = 32 - 11;   // modulus switch to Z_2N
    __shared__ int32_t s_acc[Q4_NB][2 * N];
    __shared__ float s_x[Q4_XBUF];
    const int lane = threadIdx.x, ml = lane >> 2, mb = lane & 3;
    const int g = blockIdx.x * Q4_NB + mb;
    const bool active = g < A.L.count;
    int32_t* acc = s_acc[mb];
    const int n = A.n;

    // gate inputs of this lane's bootstrap (modes 0 and 3), as boot_prologue picks them
    const int32_t* xa = nullptr;
    const int32_t* xb = nullptr;
    uint32_t al = 0, be = 0;
    int32_t cst = 0;
    if (active && (A.L.mode == 0 || A.L.mode == 3)) {
        if (A.L.mode == 0) {
            const int s1 = (A.L.nseg > 1 && g >= A.L.seg_start[1]) ? 1 : 0;
            const int gl = g - A.L.seg_start[s1];
            xa = A.L.in_a[s1] + (size_t)gl * (n + 1);
            xb = A.L.in_b[s1] ? A.L.in_b[s1] + (size_t)gl * (n + 1) : nullptr;
            al = (uint32_t)A.L.alpha[s1];
            be = (uint32_t)A.L.beta[s1];
            cst = A.L.cst[s1];
        } else {
            const int s = g / A.L.seg_jobs;
            const int gl = g - s * A.L.seg_jobs;
            const SegDesc D = A.L.segs[s];
            xa = D.a + (size_t)gl * (n + 1);
            xb = D.b ? D.b + (size_t)gl * (n + 1) : nullptr;
            al = (uint32_t)D.alpha;
            be = (uint32_t)D.beta;
            cst = D.cst;
        }
    }
    auto bara_of = [&](int i) -> int {
        if (!active) return 0;
        if (A.L.mode == 1) return A.L.bara_in[(size_t)g * n + i];
        uint32_t x = al * (uint32_t)xa[i];
        if (xb) x += be * (uint32_t)xb[i];
        return (int)((x + (1u << (SH - 1))) >> SH);
    };
    // prologue: test vector (modes 0, 3) or the accumulator (mode 1), 16 lanes per bootstrap
    if (A.L.mode == 1) {
        for (int j = ml; j < 2 * N; j += 16) acc[j] = active ? A.L.acc_io[(size_t)g * 2 * N + j] : 0;
    } else {
        int sft = 0;
        if (active) {
            uint32_t xbv = (uint32_t)cst + al * (uint32_t)xa[n];
            if (xb) xbv += be * (uint32_t)xb[n];
            const int barb = (int)((xbv + (1u << (SH - 1))) >> SH);
            sft = (2 * N - barb) & (2 * N - 1);
        }
        for (int j = ml; j < N; j += 16) {
            acc[j] = 0;
            const int e = (j - sft) & (2 * N - 1);
            acc[N + j] = active ? ((e < N) ? A.L.mu : -A.L.mu) : 0;
        }
    }
    wave_sync();

    const C2<float>* tt = (const C2<float>*)A.t1 + ml * 32;
    const __amdgpu_buffer_rsrc_t key = key_rsrc(A.bk, (uint32_t)n * Q4_ROWSET_B);
    const int sb = A.split_bits;
    int a_next = n > 0 ? bara_of(0) : 0;
    for (int i = 0; i < n; ++i) {
        const int a = a_next;
        if (i + 1 < n) a_next = bara_of(i + 1);
        q4v4f D[32][2];
        const uint32_t base4 = rot_base4(ml, a);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            // (X^a - 1) acc_q + offset, xor offset, for j = 16 t + l (t < 64).  The rotation's addresses
            // and signs are the same for both polynomials; an opaque copy of the base per polynomial
            // keeps the compiler from computing them once and holding all 64 across the transforms.
            uint32_t b4 = base4;
            asm volatile("" : "+v"(b4));
            uint32_t u[64];
#pragma unroll
            for (int t = 0; t < 64; ++t) u[t] = rotdiff_e4<N>(acc + q * N, b4 + 64u * t, 16 * t + ml, OFF) ^ OFF;
#pragma unroll
            for (int lvl = 0; lvl < L; ++lvl) {
                C2<float> v[32];
#pragma unroll
                for (int r = 0; r < 32; ++r)
                    v[r] = C2<float>{(float)digit_sx<BG>(u[r], lvl), (float)digit_sx<BG>(u[r + 32], lvl)};
                q4_fwd(v, tt, s_x, lane);
                const uint32_t ro = (uint32_t)i * Q4_ROWSET_B + (uint32_t)(q * L + lvl) * (32u * 1024u);
#pragma unroll
                for (int reg = 0; reg < 32; ++reg) {
                    // at most eight key registers in flight (the scheduler would otherwise issue all
                    // 32 loads at once: 128 registers)
                    if (reg % 8 == 0) __builtin_amdgcn_sched_barrier(0);
                    const float4 kk = key_load(key, (uint32_t)lane * 16u, ro + (uint32_t)reg * 1024u);
                    if (q == 0 && lvl == 0) {
                        const q4v4f z = {0.f, 0.f, 0.f, 0.f};
                        D[reg][0] = __builtin_amdgcn_mfma_f32_4x4x1f32(kk.x, v[reg].x, z, 0, 0, 0);
                        D[reg][1] = __builtin_amdgcn_mfma_f32_4x4x1f32(kk.z, v[reg].x, z, 0, 0, 0);
                    } else {
                        D[reg][0] = __builtin_amdgcn_mfma_f32_4x4x1f32(kk.x, v[reg].x, D[reg][0], 0, 0, 0);
                        D[reg][1] = __builtin_amdgcn_mfma_f32_4x4x1f32(kk.z, v[reg].x, D[reg][1], 0, 0, 0);
                    }
                    D[reg][0] = __builtin_amdgcn_mfma_f32_4x4x1f32(kk.y, v[reg].y, D[reg][0], 0, 0, 0);
                    D[reg][1] = __builtin_amdgcn_mfma_f32_4x4x1f32(kk.w, v[reg].y, D[reg][1], 0, 0, 0);
                }
            }
        }
        // inverse transforms; FP64 untwist and rounding; acc += (hi << split) + lo
#pragma unroll
        for (int o = 0; o < 2; ++o) {
            uint32_t rh[64];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                C2<float> v[32];
#pragma unroll
                for (int reg = 0; reg < 32; ++reg) v[reg] = C2<float>{D[reg][o][2 * h], D[reg][o][2 * h + 1]};
                q4_inv(v, tt, s_x, lane);
#pragma unroll
                for (int r = 0; r < 32; ++r) {
                    const C2<float> y = v[q4_slot(r)];
                    const C2<double> z = cmulc(C2<double>{(double)y.x, (double)y.y}, C2<double>{kQ4UntwRe[r], kQ4UntwIm[r]});
                    if (h == 0) {
                        rh[2 * r] = round_mod32(z.x);
                        rh[2 * r + 1] = round_mod32(z.y);
                    } else {
                        const int j = o * N + 16 * r + ml;
                        acc[j] = (int32_t)((uint32_t)acc[j] + (rh[2 * r] << sb) + round_mod32(z.x));
                        acc[j + 512] = (int32_t)((uint32_t)acc[j + 512] + (rh[2 * r + 1] << sb) + round_mod32(z.y));
                    }
                }
            }
        }
        wave_sync();
    }
    // epilogue: extraction (modes 0, 3) or the accumulator (mode 1)
    if (active) {
        if (A.L.mode == 1) {
            for (int j = ml; j < 2 * N; j += 16) A.L.acc_io[(size_t)g * 2 * N + j] = acc[j];
        } else {
            int32_t* u = A.L.u_out + (size_t)g * (N + 1);
            for (int j = ml; j < N; j += 16) u[j] = (j == 0) ? acc[0] : (int32_t)(0u - (uint32_t)acc[N - j]);
            if (ml == 0) u[N] = acc[N];
            if (A.L.acc_io)   // the accumulator too, when the launch asks for it (as boot_epilogue)
                for (int j = ml; j < 2 * N; j += 16) A.L.acc_io[(size_t)g * 2 * N + j] = acc[j];
        }
    }
}

// The key of k_boot32q from the split key bk32 ([n][2l][2][512] float4 {hi.re, hi.im, lo.re,
// lo.im}, frequency f natural): element (i, row, register 16 s + k2, lane 4 l' + c) holds component c
// of A_re(K_0), A_im(K_0), A_re(K_1), A_im(K_1) at f = l' + 16 s + 32 k2, where A_re(K) = (hr, hi, lr,
// li) and A_im(K) = (-hi, hr, -li, lr).  One thread per element.
__global__ void __launch_bounds__(256) k_keyprep_q4(const float4* __restrict__ bk32, float4* __restrict__ out,
                                                    size_t total) {
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= total) return;
    const int lane = (int)(e & 63), reg = (int)((e >> 6) & 31);
    const size_t rowi = e >> 11;   // (i * 6 + row)
    const int c = lane & 3, lp = lane >> 2;
    const int f = lp + 16 * (reg >> 4) + 32 * (reg & 15);
    float a[2][2];
#pragma unroll
    for (int o = 0; o < 2; ++o) {
        const float4 K = bk32[(rowi * 2 + o) * 512 + f];
        const float re[4] = {K.x, K.y, K.z, K.w};
        const float im[4] = {-K.y, K.x, -K.w, K.z};
        a[o][0] = re[c];
        a[o][1] = im[c];
    }
    out[e] = make_float4(a[0][0], a[0][1], a[1][0], a[1][1]);
}

// Transform test of the 4-per-wavefront layout: forward (int32 [count][1024] -> F natural order,
// interleaved doubles [count][1024]) or inverse (F -> integer-unit doubles [count][1024]).
__global__ void __launch_bounds__(64) k_fft_test_q4(double* __restrict__ out, const void* __restrict__ in,
                                                    int inverse, int count, const C2<float>* __restrict__ q4tw) {
    constexpr int N = 1024, M = 512;
    __shared__ float s_x[Q4_XBUF];
    const int lane = threadIdx.x, ml = lane >> 2, mb = lane & 3;
    const int g = blockIdx.x * Q4_NB + mb;
    const bool active = g < count;
    const C2<float>* tt = q4tw + ml * 32;
    C2<float> v[32];
    if (!inverse) {
        const int32_t* p = (const int32_t*)in + (size_t)(active ? g : 0) * N;
#pragma unroll
        for (int r = 0; r < 32; ++r)
            v[r] = active ? C2<float>{(float)p[16 * r + ml], (float)p[16 * r + ml + M]} : C2<float>{0.f, 0.f};
        q4_fwd(v, tt, s_x, lane);
        if (active) {
            double* o = out + (size_t)g * N;
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int k2 = 0; k2 < 16; ++k2) {
                    const int f = ml + 16 * s + 32 * k2;
                    o[2 * f] = v[16 * s + k2].x;
                    o[2 * f + 1] = v[16 * s + k2].y;
                }
        }
    } else {
        const double* F = (const double*)in + (size_t)(active ? g : 0) * N;
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int k2 = 0; k2 < 16; ++k2) {
                const int f = ml + 16 * s + 32 * k2;
                v[16 * s + k2] = active ? C2<float>{(float)F[2 * f], (float)F[2 * f + 1]} : C2<float>{0.f, 0.f};
            }
        q4_inv(v, tt, s_x, lane);
        if (active) {
            double* o = out + (size_t)g * N;
#pragma unroll
            for (int r = 0; r < 32; ++r) {
                const C2<float> y = v[q4_slot(r)];
                const C2<double> z = cmulc(C2<double>{(double)y.x, (double)y.y}, C2<double>{kQ4UntwRe[r], kQ4UntwIm[r]});
                o[16 * r + ml] = z.x;
                o[16 * r + ml + M] = z.y;
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Fixed scratch for the entry point that allocates nothing (tfhe_gpu_gate_batch_dev_noalloc): the
// extracted samples between the bootstrap and the keyswitch live in this array instead of a buffer
// grown with hipMalloc.  Host code never uses its name as an address: scratch_init() resolves the
// device address once per device (with that device current) and every launch takes that pointer.
// ---------------------------------------------------------------------------------------------
static __device__ __align__(256) int32_t scratch[kScratchElems];
static void* g_scratch_ptr[kScratchMaxDevices] = {};

hipError_t scratch_init(int device) {
    if (device < 0 || device >= kScratchMaxDevices) return hipErrorInvalidDevice;
    if (g_scratch_ptr[device]) return hipSuccess;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = p;
    return hipSuccess;
}

int32_t* scratch_ptr(int device) {
    return (device >= 0 && device < kScratchMaxDevices) ? (int32_t*)g_scratch_ptr[device] : nullptr;
}

// FP64 bootstrap form when a launch does not name one: 0 one bootstrap per workgroup, 1 workgroups
// sharing key rows, 2 (default) by launch size (boot64_single_round, boot64_latency_max)
static int g_boot64_shared = 2;
void set_boot64_shared(int mode) { g_boot64_shared = mode; }

// split-FP32 MAC form at N = 1024: 0 plain FMA accumulation (default), 1 compensated
static int g_split_mac = 0;
void set_split_mac(int mode) { g_split_mac = mode; }

// launch form applied to launches that name none (0: the size rules); see BootLaunch::form
static int g_forced_form = 0;
void set_forced_form(int form) { g_forced_form = form; }
int forced_form() { return g_forced_form; }

// keyswitch fused into the split-FP32 bootstrap kernel (an evaluated alternative; off by default)
static int g_fused_ks = 0;
void set_fused_keyswitch(int on) { g_fused_ks = on; }
int fused_keyswitch() { return g_fused_ks; }

static int device_cus() {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
    }
    return cus;
}

// Launches of at most this many bootstraps take the latency form (k_boot64lat): with its 133 KB of
// LDS one workgroup runs per CU, and one round of its short steps beats the one round of the
// one-per-workgroup form (4.4 against 7.0 ms for a gate); two rounds do not (8.5 against 8.0 ms
// for 512 gates on 256 CUs).
static int boot64_latency_max() { return device_cus(); }

// Bootstraps the GPU runs at once in the one-per-workgroup FP64 form: its 35 KB of LDS fits four
// workgroups per CU.  A launch of at most that many finishes in one round of that form, which
// beats the key-sharing form's round (same step count, two wavefronts per SIMD in lockstep).
static int boot64_single_round() { return 4 * device_cus(); }

template <int N, int L, int BG>
static hipError_t launch_boot_t(const DevKeyView& key, const DevTwiddles& tw, Prec prec, const BootLaunch& Lc,
                                hipStream_t st) {
    BootKArgs A{};
    A.L = Lc;
    A.n = key.p.n;
    A.split_bits = key.split_bits;
    A.wa64 = tw.wa64;
    A.twist32 = tw.twist32;
    A.untw64 = tw.untw64;
    A.fwd_a16 = tw.fwd_a16;
    A.fwd_t2 = tw.fwd_t2;
    if (prec == PREC_FP64) {
        A.bk = key.bk64;
        A.t1 = tw.t1_fold64;
        A.t2 = tw.t2_64;
        if constexpr (N == 1024) {
            int form = Lc.form ? Lc.form : g_forced_form;
            if (form > 3) form = 0;   // the N = 2048 workgroup sizes: the size rules here
            if (Lc.mode != 2 && key.p.n <= kMaxLweN &&
                (form == 3 || (form == 0 && g_boot64_shared == 2 && Lc.count <= boot64_latency_max()))) {
                hipLaunchKernelGGL((k_boot64lat<N, L, BG>), dim3(Lc.count), dim3(64 * 2 * L), 0, st, A);
                return hipGetLastError();
            }
            const bool shared = form == 1 ||
                                (form == 0 && (g_boot64_shared == 1 ||
                                               (g_boot64_shared == 2 && Lc.count > boot64_single_round())));
            if (Lc.mode != 2 && key.p.n <= B64W_MAXN && shared) {
                const int blocks = (Lc.count + B64W_GATES - 1) / B64W_GATES;
                hipLaunchKernelGGL((k_boot64w<N, L, BG>), dim3(blocks), dim3(64 * B64W_GATES), 0, st, A);
                return hipGetLastError();
            }
        }
        hipLaunchKernelGGL((k_boot64<N, L, BG>), dim3(Lc.count), dim3(64), 0, st, A);
    } else {
        A.bk = key.bk32;
        A.t1 = tw.t1_plain32;   // W_512^{c L}: pass-3 input twiddles of the decimation-in-time transform
        A.t2 = tw.t2dit32;
        if constexpr (N == 1024) {
            if (g_split_mac == 1) {
                // the compensated-accumulation form exists for the one-bootstrap-per-wavefront kernel
                hipLaunchKernelGGL((k_boot32<N, L, BG, 1>), dim3(Lc.count), dim3(64), 0, st, A);
                return hipGetLastError();
            }
            if (Lc.ks_out) {
                // fused keyswitch: mode 0, one gate segment, the (t, NP) = (8, 640) keyswitch key
                if (Lc.mode != 0 || Lc.nseg != 1 || key.ksk_np != 640 || key.p.ks_t != 8 || !key.ksk)
                    return hipErrorInvalidValue;
                A.ksk = key.ksk;
                hipLaunchKernelGGL((k_boot32<N, L, BG, 0, true>), dim3(Lc.count), dim3(64), 0, st, A);
                return hipGetLastError();
            }
            // forms: 6 four bootstraps per wavefront with the products on the matrix cores, 3 the
            // latency form, 2 one bootstrap per wavefront, 0 (and the forms of other paths) by
            // launch size
            int form = Lc.form ? Lc.form : g_forced_form;
            if (form == 6 && Lc.mode != 2 && key.bk32q && key.p.l == 3 && key.p.n <= kMaxLweN) {
                A.bk = key.bk32q;
                A.t1 = tw.q4tw;
                hipLaunchKernelGGL((k_boot32q<L, BG>), dim3((Lc.count + Q4_NB - 1) / Q4_NB), dim3(64), 0, st, A);
                return hipGetLastError();
            }
            if (form != 2 && form != 3) form = 0;
            if (Lc.mode != 2 && (form == 3 || (form == 0 && Lc.count <= boot64_latency_max()))) {
                hipLaunchKernelGGL((k_boot32lat<N, L, BG>), dim3(Lc.count), dim3(64 * 2 * L), 0, st, A);
                return hipGetLastError();
            }
        }
        hipLaunchKernelGGL((k_boot32<N, L, BG>), dim3(Lc.count), dim3(64), 0, st, A);
    }
    return hipGetLastError();
}

hipError_t launch_bootstrap(const DevKeyView& key, const DevTwiddles& tw, Prec prec, const BootLaunch& Lc,
                            hipStream_t st) {
    if (Lc.count <= 0) return hipSuccess;
    if ((prec == PREC_FP64 && !key.bk64) || (prec == PREC_SPLIT32 && !key.bk32)) return hipErrorInvalidValue;
    if (key.p.N == 1024) return launch_boot_t<1024, 3, 7>(key, tw, prec, Lc, st);
    if (key.p.N == 2048) {
        // two wavefronts per bootstrap (k_boot2w); every mode is supported
        BootKArgs A{};
        A.L = Lc;
        A.n = key.p.n;
        A.split_bits = key.split_bits;
        A.t2w_rw64 = tw.rw64;
        A.t2w_rw32 = tw.rw32;
        A.fwd_a16 = tw.w2_a16;
        A.fwd_t2 = tw.w2_t1p32;
        // launches of at most one bootstrap per CU take the latency form (FP64: unless a launch form
        // other than it is forced, as at N = 1024).  Forms: 3 latency; 2, 4, 5 the batched kernel
        // with 1, 2, 4 bootstraps per workgroup; 0 by launch size (2 per workgroup).  Two per
        // workgroup, two workgroups per CU, beat one (the rows' reads no longer meet in L1) and four
        // (a longer barrier lockstep): per step of 8192 MUX + 8192 XOR on one box, FP64 652 / 482 /
        // 502 ms and split 331 / 306 / 328 ms for 1 / 2 / 4.
        const int form = Lc.form ? Lc.form : g_forced_form;
        const bool lat = Lc.mode != 2 && A.n <= kMaxLweN &&
                         (form == 3 || (form == 0 && (prec == PREC_SPLIT32 || g_boot64_shared == 2) &&
                                        Lc.count <= boot64_latency_max()));
        const int gb = form == 2 ? 1 : form == 5 ? 4 : 2;
        if (prec == PREC_FP64) {
            A.bk = key.bk64;
            A.t1 = tw.t1w64;
            A.t2 = tw.t2_64;
            if (lat) {
                hipLaunchKernelGGL((k_boot2wlat<3, 8, 0>), dim3(Lc.count), dim3(64 * 6), 0, st, A);
                return hipGetLastError();
            }
            if (gb == 1) hipLaunchKernelGGL((k_boot2w<3, 8, 0, 1>), dim3(Lc.count), dim3(128), 0, st, A);
            else if (gb == 2) hipLaunchKernelGGL((k_boot2w<3, 8, 0, 2>), dim3((Lc.count + 1) / 2), dim3(256), 0, st, A);
            else hipLaunchKernelGGL((k_boot2w<3, 8, 0, 4>), dim3((Lc.count + 3) / 4), dim3(512), 0, st, A);
        } else {
            A.bk = key.bk32;
            A.t1 = tw.t1w32;
            A.t2 = tw.t2_32;
            A.untw64 = tw.rws64;
            if (lat) {
                hipLaunchKernelGGL((k_boot2wlat<3, 8, 1>), dim3(Lc.count), dim3(64 * 8), 0, st, A);
                return hipGetLastError();
            }
            if (gb == 1) hipLaunchKernelGGL((k_boot2w<3, 8, 1, 1>), dim3(Lc.count), dim3(128), 0, st, A);
            else if (gb == 4) hipLaunchKernelGGL((k_boot2w<3, 8, 1, 4>), dim3((Lc.count + 3) / 4), dim3(512), 0, st, A);
            else hipLaunchKernelGGL((k_boot2w<3, 8, 1, 2>), dim3((Lc.count + 1) / 2), dim3(256), 0, st, A);
        }
        return hipGetLastError();
    }
    return hipErrorInvalidValue;
}

template <int TKS, int KN, int NP>
static hipError_t launch_ks_t(const DevKeyView& key, int32_t* out, const int32_t* u1, const int32_t* u2,
                              int32_t add_b, int count, hipStream_t st, int32_t* const* out_tab, int out_seg) {
    constexpr int C = TFHE_KS_C;
    if (key.ksk_np != NP || key.p.N != KN) return hipErrorInvalidValue;
    if (count <= KS_SMALL_MAX) {
        hipLaunchKernelGGL(k_ks_clear, dim3(count), dim3(256), 0, st, out, key.p.n, out_tab, out_seg);
        hipLaunchKernelGGL((k_keyswitch_sliced<TKS, KN, NP>), dim3(NP / 64, KS_SLICES, count), dim3(64), 0, st, u1,
                           u2, add_b, key.ksk, out, key.p.n, out_tab, out_seg);
        return hipGetLastError();
    }
    dim3 grid(NP / C, (count + KS_GATES - 1) / KS_GATES);
    hipLaunchKernelGGL((k_keyswitch<TKS, KN, NP, C>), grid, dim3(KS_GATES), 0, st, u1, u2, add_b, key.ksk, out, count,
                       key.p.n, out_tab, out_seg);
    return hipGetLastError();
}

hipError_t launch_keyswitch(const DevKeyView& key, int32_t* out, const int32_t* u1, const int32_t* u2, int32_t add_b,
                            int count, hipStream_t st, int32_t* const* out_tab, int out_seg) {
    if (count <= 0) return hipSuccess;
    if (out_tab && out_seg <= 0) return hipErrorInvalidValue;
    // the two supported parameter sets: (t, N, padded n+1) = (8, 1024, 640) and (10, 2048, 768)
    if (key.p.ks_t == 8 && key.p.N == 1024 && key.ksk_np == 640)
        return launch_ks_t<8, 1024, 640>(key, out, u1, u2, add_b, count, st, out_tab, out_seg);
    if (key.p.ks_t == 10 && key.p.N == 2048 && key.ksk_np == 768)
        return launch_ks_t<10, 2048, 768>(key, out, u1, u2, add_b, count, st, out_tab, out_seg);
    return hipErrorInvalidValue;
}

hipError_t launch_keyprep(const DevParams& p, const DevTwiddles& tw, const int32_t* d_bk, void* bk64, void* bk32,
                          int split_bits, hipStream_t st) {
    const int npoly = p.n * 2 * p.l * 2;
    if (p.N == 1024)
        hipLaunchKernelGGL((k_keyprep<1024>), dim3(npoly), dim3(64), 0, st, d_bk, (C2<double>*)bk64, (float4*)bk32,
                           split_bits, (const C2<double>*)tw.t1_fold64, (const C2<double>*)tw.t2_64,
                           (const C2<double>*)tw.wa64);
    else if (p.N == 2048)
        hipLaunchKernelGGL(k_keyprep2w, dim3(npoly), dim3(128), 0, st, d_bk, (C2<double>*)bk64, (float4*)bk32,
                           split_bits, (const C2<double>*)tw.t1w64, (const C2<double>*)tw.t2_64,
                           (const C2<double>*)tw.rw64);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_keyprep_q4(const DevParams& p, const void* bk32, void* bk32q, hipStream_t st) {
    if (p.N != 1024 || p.l != 3) return hipErrorInvalidValue;
    const size_t total = (size_t)p.n * 6 * 32 * 64;
    hipLaunchKernelGGL(k_keyprep_q4, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, (const float4*)bk32,
                       (float4*)bk32q, total);
    return hipGetLastError();
}

hipError_t launch_fft_test_q4(const DevTwiddles& tw, int inverse, double* out, const void* in, int count,
                              hipStream_t st) {
    if (count <= 0) return hipSuccess;
    if (!tw.q4tw) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_fft_test_q4, dim3((count + Q4_NB - 1) / Q4_NB), dim3(64), 0, st, out, in, inverse, count,
                       (const C2<float>*)tw.q4tw);
    return hipGetLastError();
}

hipError_t launch_ksk_relayout(const DevParams& p, const int32_t* in, int32_t* out, int np, hipStream_t st) {
    hipLaunchKernelGGL(k_ksk_relayout, dim3(4096), dim3(256), 0, st, in, out, p.N, p.ks_t, p.n, np);
    return hipGetLastError();
}

hipError_t launch_extract(int N, int32_t* u, const int32_t* acc, int count, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    if (N == 1024) hipLaunchKernelGGL((k_extract<1024>), dim3(count), dim3(256), 0, st, u, acc, count);
    else if (N == 2048) hipLaunchKernelGGL((k_extract<2048>), dim3(count), dim3(256), 0, st, u, acc, count);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_rotate_decompose(const DevParams& p, int32_t* rot, int32_t* digits, const int32_t* acc,
                                   const int32_t* bara, int count, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    if (p.N == 1024) hipLaunchKernelGGL((k_rotdec<1024, 3, 7>), dim3(count), dim3(64), 0, st, rot, digits, acc, bara);
    else if (p.N == 2048) hipLaunchKernelGGL((k_rotdec<2048, 3, 8>), dim3(count), dim3(64), 0, st, rot, digits, acc, bara);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_fft_test(const DevTwiddles& tw, int N, Prec prec, int inverse, double* out, const void* in, int count,
                           hipStream_t st) {
    if (count <= 0) return hipSuccess;
    if (prec == PREC_FP64) {
        if (N == 1024)
            hipLaunchKernelGGL((k_fft_test64<1024>), dim3(count), dim3(64), 0, st, out, in, inverse,
                               (const C2<double>*)tw.t1_fold64, (const C2<double>*)tw.t2_64, (const C2<double>*)tw.wa64);
        else if (N == 2048)
            hipLaunchKernelGGL((k_fft_test2w<0>), dim3(count), dim3(128), 0, st, out, in, inverse, tw.t1w64, tw.t2_64,
                               (const C2<float>*)tw.rw32, (const C2<double>*)tw.rw64, (const C2<double>*)tw.rws64, nullptr, nullptr);
        else return hipErrorInvalidValue;
    } else {
        if (N == 1024)
            hipLaunchKernelGGL((k_fft_test32<1024>), dim3(count), dim3(64), 0, st, out, in, inverse,
                               (const C2<float>*)tw.t1_plain32, (const C2<float>*)tw.t2dit32,
                               (const C2<float>*)tw.twist32, (const C2<double>*)tw.untw64,
                               (const f16x8_t*)tw.fwd_a16, (const C2<float>*)tw.fwd_t2);
        else if (N == 2048)
            hipLaunchKernelGGL((k_fft_test2w<1>), dim3(count), dim3(128), 0, st, out, in, inverse, tw.t1w32, tw.t2_32,
                               (const C2<float>*)tw.rw32, (const C2<double>*)tw.rw64, (const C2<double>*)tw.rws64, tw.w2_a16, tw.w2_t1p32);
        else return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// The split kernels' digit conversion on its own: pair i is (x[2i], x[2i + 1]), taken as the
// rotation leaves them (x + offset, no xor); out[(i * 3 + lvl) * 2 + h] = digit lvl of x[2i + h].
template <int BG>
__global__ void __launch_bounds__(256) k_digits_test(float* __restrict__ out, const int32_t* __restrict__ x,
                                                     int variant, int count) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    constexpr uint32_t OFF = decomp_offset<3, BG>();
    const uint32_t a = (uint32_t)x[2 * i] + OFF, b = (uint32_t)x[2 * i + 1] + OFF;
    uint32_t p0 = 0, p1 = 0;
    if constexpr (BG == 7) digits7_prep(a, b, p0, p1);
#pragma unroll
    for (int lvl = 0; lvl < 3; ++lvl) {
        uint32_t w;
        if constexpr (BG == 7) w = variant ? digits7_f16x2_one(a, b, lvl) : digits7_f16x2(p0, p1, lvl);
        else w = digits8_f16x2(a, b, lvl);
        const f16x2_t h = __builtin_bit_cast(f16x2_t, w);
        out[(i * 3 + lvl) * 2] = (float)h[0];
        out[(i * 3 + lvl) * 2 + 1] = (float)h[1];
    }
}

hipError_t launch_digits_test(int bg, int variant, float* out, const int32_t* x, int count, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    const dim3 grid((count + 255) / 256);
    if (bg == 7) hipLaunchKernelGGL(k_digits_test<7>, grid, dim3(256), 0, st, out, x, variant, count);
    else if (bg == 8) hipLaunchKernelGGL(k_digits_test<8>, grid, dim3(256), 0, st, out, x, variant, count);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_valu_probe(int fp64, void* out, int blocks, int iters, hipStream_t st) {
    if (fp64)
        hipLaunchKernelGGL((k_valu_probe<double>), dim3(blocks), dim3(256), 0, st, (double*)out, 1.0000001, 1e-9, iters);
    else
        hipLaunchKernelGGL((k_valu_probe<float>), dim3(blocks), dim3(256), 0, st, (float*)out, 1.0000001f, 1e-9f, iters);
    return hipGetLastError();
}

hipError_t launch_decrypt(int32_t* bits, const int32_t* samples, const int32_t* key, int n, int count, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_decrypt, dim3(count), dim3(64), 0, st, bits, samples, key, n);
    return hipGetLastError();
}

}  // namespace tfhe
