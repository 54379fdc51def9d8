// tfhe_device.hpp -- device building blocks of the gfx950 bootstrap engine.
//
// Folded negacyclic transform.  A real polynomial p of degree < N (mod X^N + 1) is represented
// by its values at the M = N/2 points w^(4k+1), w = e^{i pi / N}:
//     z_j = (p_j + i p_{j+M}) w^j,      F_k = sum_{j<M} z_j e^{+2 pi i jk / M}.
// These are the odd powers of w that the reference's 2N-point embedding evaluates
// (kernel/src/tfhe_blindRotate_FFT/tGswFFTExternMulToTLwe.cpp:97-130), up to an exact
// permutation and conjugation, at a quarter of the transform size.
//
// One 64-lane wavefront owns one transform of M = 64 * R1 points (R1 = 8 for N = 1024,
// 16 for N = 2048).  Index split j = 64 a + 8 b + c and k = k1 + R1 k2 + 8 R1 k3 gives three
// register-resident passes (radix R1 over a, radix 8 over b, radix 8 over c) with two LDS
// exchanges between them:
//   pass 1  lane (b,c) = 8b+c, registers a -> k1, then twiddle W_M^{(8b+c) k1}
//   pass 2  lane (c,k1) = c + 8(k1 mod 8), registers b -> k2, then twiddle W_64^{c k2}
//   pass 3  lane (k1,k2) = (k1 mod 8) + 8 k2, registers c -> k3
// The output stays in this "lane order" (for R1 = 8 it is natural order: lane L, register r
// holds F_{L + 64 r}); the key is stored in the same order, so the pointwise product needs no
// permutation, and the inverse runs the passes backwards from lane order to natural j order.
// The split-FP32 kernel uses the decimation-in-time form of the same factorisation instead
// (fft512_dit: twiddles on the butterfly inputs, fused into them, one code for both directions).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tfhe {

template <typename T>
struct C2 {
    T x, y;
};

template <typename T>
__device__ __forceinline__ C2<T> operator+(C2<T> a, C2<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ C2<T> operator-(C2<T> a, C2<T> b) { return {a.x - b.x, a.y - b.y}; }

// a * b
template <typename T>
__device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b) {
    return {__builtin_fma(a.x, b.x, -(a.y * b.y)), __builtin_fma(a.x, b.y, a.y * b.x)};
}
__device__ __forceinline__ C2<float> cmul(C2<float> a, C2<float> b) {
    return {__builtin_fmaf(a.x, b.x, -(a.y * b.y)), __builtin_fmaf(a.x, b.y, a.y * b.x)};
}
// a * conj(b)
template <typename T>
__device__ __forceinline__ C2<T> cmulc(C2<T> a, C2<T> b) {
    return {__builtin_fma(a.x, b.x, a.y * b.y), __builtin_fma(a.y, b.x, -(a.x * b.y))};
}
__device__ __forceinline__ C2<float> cmulc(C2<float> a, C2<float> b) {
    return {__builtin_fmaf(a.x, b.x, a.y * b.y), __builtin_fmaf(a.y, b.x, -(a.x * b.y))};
}
// acc += a * b
template <typename T>
__device__ __forceinline__ void cmac(C2<T>& acc, C2<T> a, C2<T> b) {
    acc.x = __builtin_fma(a.x, b.x, acc.x);
    acc.x = __builtin_fma(-a.y, b.y, acc.x);
    acc.y = __builtin_fma(a.x, b.y, acc.y);
    acc.y = __builtin_fma(a.y, b.x, acc.y);
}
__device__ __forceinline__ void cmac(C2<float>& acc, C2<float> a, C2<float> b) {
    acc.x = __builtin_fmaf(a.x, b.x, acc.x);
    acc.x = __builtin_fmaf(-a.y, b.y, acc.x);
    acc.y = __builtin_fmaf(a.x, b.y, acc.y);
    acc.y = __builtin_fmaf(a.y, b.x, acc.y);
}

// s + e = a + b exactly (Knuth's TwoSum: no ordering of |a| and |b| assumed)
__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
    s = a + b;
    const float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// acc += a * b as a compensated dot-product step (the Dot2 scheme of Ogita, Rump and Oishi, 2005):
// each real product is split into its rounded value and its exact rounding error (one FMA), the
// rounded values are added to acc with TwoSum, and every rounding error -- of the products and of
// the sums -- is gathered in comp.  acc + comp then carries the dot product as if it were computed
// in twice the working precision.  About 19 operations per real component against 2 FMAs.
__device__ __forceinline__ void cmac_comp(C2<float>& acc, C2<float>& comp, C2<float> a, C2<float> b) {
    const float p0 = a.x * b.x, e0 = __builtin_fmaf(a.x, b.x, -p0);
    const float p1 = -a.y * b.y, e1 = __builtin_fmaf(-a.y, b.y, -p1);
    const float p2 = a.x * b.y, e2 = __builtin_fmaf(a.x, b.y, -p2);
    const float p3 = a.y * b.x, e3 = __builtin_fmaf(a.y, b.x, -p3);
    float t0, t1, t2, t3;
    two_sum(acc.x, p0, acc.x, t0);
    two_sum(acc.x, p1, acc.x, t1);
    two_sum(acc.y, p2, acc.y, t2);
    two_sum(acc.y, p3, acc.y, t3);
    comp.x += (e0 + e1) + (t0 + t1);
    comp.y += (e2 + e3) + (t2 + t3);
}

// multiply by e^{S i pi/2}: S=+1 -> i, S=-1 -> -i
template <int S, typename T>
__device__ __forceinline__ C2<T> mul_i(C2<T> a) {
    if constexpr (S > 0) return {-a.y, a.x};
    else return {a.y, -a.x};
}
// multiply by e^{S i pi/4}
template <int S, typename T>
__device__ __forceinline__ C2<T> mul_w8(C2<T> a) {
    const T r = (T)0.70710678118654752440084436210485;
    if constexpr (S > 0) return {(a.x - a.y) * r, (a.x + a.y) * r};
    else return {(a.x + a.y) * r, (a.y - a.x) * r};
}
// multiply by e^{3 S i pi/4}
template <int S, typename T>
__device__ __forceinline__ C2<T> mul_w8_3(C2<T> a) {
    const T r = (T)0.70710678118654752440084436210485;
    if constexpr (S > 0) return {(-a.x - a.y) * r, (a.x - a.y) * r};
    else return {(a.y - a.x) * r, (-a.x - a.y) * r};
}

// b + r q for a real r
template <typename T>
__device__ __forceinline__ C2<T> caxpy(T r, C2<T> q, C2<T> b) {
    return {__builtin_fma(r, q.x, b.x), __builtin_fma(r, q.y, b.y)};
}
__device__ __forceinline__ C2<float> caxpy(float r, C2<float> q, C2<float> b) {
    return {__builtin_fmaf(r, q.x, b.x), __builtin_fmaf(r, q.y, b.y)};
}

// Stages 2 and 3 of the 8-point DFT below (FUSE form), from the first-stage sums and differences
// a_n = v_n + v_{n+4} (n < 4), a4 = v0 - v4, d5 = v1 - v5, a6 = e^{S i pi/2} (v2 - v6), d7 = v3 - v7.
template <int S, typename T>
__device__ __forceinline__ void dft8_tail(C2<T>* v, C2<T> a0, C2<T> a4, C2<T> a1, C2<T> d5, C2<T> a2, C2<T> a6,
                                          C2<T> a3, C2<T> d7) {
    const T r = (T)0.70710678118654752440084436210485;
    // e^{S i pi/4} d5 = r p5, e^{3 S i pi/4} d7 = r p7
    const C2<T> p5 = S > 0 ? C2<T>{d5.x - d5.y, d5.x + d5.y} : C2<T>{d5.x + d5.y, d5.y - d5.x};
    const C2<T> p7 = S > 0 ? C2<T>{-d7.x - d7.y, d7.x - d7.y} : C2<T>{d7.y - d7.x, -d7.x - d7.y};
    C2<T> b0 = a0 + a2, b2 = a0 - a2, b1 = a1 + a3, b3 = mul_i<S>(a1 - a3);
    C2<T> b4 = a4 + a6, b6 = a4 - a6, q5 = p5 + p7, q7 = mul_i<S>(p5 - p7);
    v[0] = b0 + b1; v[4] = b0 - b1;
    v[2] = b2 + b3; v[6] = b2 - b3;
    v[1] = caxpy(r, q5, b4); v[5] = caxpy(-r, q5, b4);
    v[3] = caxpy(r, q7, b6); v[7] = caxpy(-r, q7, b6);
}

// In-place 8-point DFT, X_k = sum_n x_n e^{S 2 pi i nk/8}, natural order in and out.  The odd
// half carries e^{S i pi/4} and e^{3 S i pi/4}; both share the factor 1/sqrt(2), which FUSE
// applies once, fused into the last stage (52 real operations, no plain multiplies).  Without
// FUSE the factor is applied on the way in, which keeps fewer values live (for kernels at their
// register limit).
template <int S, typename T, bool FUSE = true>
__device__ __forceinline__ void dft8(C2<T>* v) {
    if constexpr (!FUSE) {
        C2<T> a0 = v[0] + v[4], a4 = v[0] - v[4];
        C2<T> a1 = v[1] + v[5], a5 = mul_w8<S>(v[1] - v[5]);
        C2<T> a2 = v[2] + v[6], a6 = mul_i<S>(v[2] - v[6]);
        C2<T> a3 = v[3] + v[7], a7 = mul_w8_3<S>(v[3] - v[7]);
        C2<T> b0 = a0 + a2, b2 = a0 - a2, b1 = a1 + a3, b3 = mul_i<S>(a1 - a3);
        C2<T> b4 = a4 + a6, b6 = a4 - a6, b5 = a5 + a7, b7 = mul_i<S>(a5 - a7);
        v[0] = b0 + b1; v[4] = b0 - b1;
        v[2] = b2 + b3; v[6] = b2 - b3;
        v[1] = b4 + b5; v[5] = b4 - b5;
        v[3] = b6 + b7; v[7] = b6 - b7;
        return;
    }
    dft8_tail<S>(v, v[0] + v[4], v[0] - v[4], v[1] + v[5], v[1] - v[5], v[2] + v[6], mul_i<S>(v[2] - v[6]),
                 v[3] + v[7], v[3] - v[7]);
}

__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }

// a * w, or a * conj(w) when CONJ
template <bool CONJ, typename T>
__device__ __forceinline__ C2<T> cmul_w(C2<T> a, C2<T> w) {
    if constexpr (CONJ) return cmulc(a, w);
    else return cmul(a, w);
}

// Radix-2 butterfly whose second input carries a twiddle: s = x0 + x1 w, d = x0 - x1 w (conj(w)
// when CONJ).  Six FMAs, the difference taken as d = 2 x0 - s, instead of a four-instruction
// complex multiply and four adds.
template <bool CONJ, typename T>
__device__ __forceinline__ void bfly_tw(C2<T> x0, C2<T> x1, C2<T> w, C2<T>& s, C2<T>& d) {
    if constexpr (!CONJ) {
        s.x = fmaT(x1.x, w.x, fmaT(-x1.y, w.y, x0.x));
        s.y = fmaT(x1.x, w.y, fmaT(x1.y, w.x, x0.y));
    } else {
        s.x = fmaT(x1.x, w.x, fmaT(x1.y, w.y, x0.x));
        s.y = fmaT(x1.y, w.x, fmaT(-x1.x, w.y, x0.y));
    }
    d.x = fmaT((T)2, x0.x, -s.x);
    d.y = fmaT((T)2, x0.y, -s.y);
}

// The 8-point DFT (FUSE form) of the twiddled inputs v_n w_n (conj(w_n) when CONJ): each twiddle is
// fused into a first-stage butterfly (bfly_tw).  UNIT0: w_0 = 1 and w[0] is not read.  With
// UNIT0, 72 real operations against 80 for seven complex multiplies and dft8.
template <int S, bool CONJ, bool UNIT0, typename T>
__device__ __forceinline__ void dft8_tw(C2<T>* v, const C2<T>* w) {
    C2<T> a0, a4, a1, d5, a2, d6, a3, d7;
    bfly_tw<CONJ>(UNIT0 ? v[0] : cmul_w<CONJ>(v[0], w[0]), v[4], w[4], a0, a4);
    bfly_tw<CONJ>(cmul_w<CONJ>(v[1], w[1]), v[5], w[5], a1, d5);
    bfly_tw<CONJ>(cmul_w<CONJ>(v[2], w[2]), v[6], w[6], a2, d6);
    bfly_tw<CONJ>(cmul_w<CONJ>(v[3], w[3]), v[7], w[7], a3, d7);
    dft8_tail<S>(v, a0, a4, a1, d5, a2, mul_i<S>(d6), a3, d7);
}

// In-place 16-point DFT: one radix-2 DIF step, then two 8-point DFTs.
template <int S, typename T>
__device__ __forceinline__ void dft16(C2<T>* v) {
    // twiddles e^{S 2 pi i n/16}, n = 0..7
    const T c1 = (T)0.92387953251128675612818318939679, s1 = (T)0.38268343236508977172845998403040;
    const T r = (T)0.70710678118654752440084436210485;
    C2<T> e[8], o[8];
#pragma unroll
    for (int n = 0; n < 8; ++n) { e[n] = v[n] + v[n + 8]; o[n] = v[n] - v[n + 8]; }
    const T sg = (T)S;
    o[1] = cmul(o[1], C2<T>{c1, sg * s1});
    o[2] = cmul(o[2], C2<T>{r, sg * r});
    o[3] = cmul(o[3], C2<T>{s1, sg * c1});
    o[4] = mul_i<S>(o[4]);
    o[5] = cmul(o[5], C2<T>{-s1, sg * c1});
    o[6] = cmul(o[6], C2<T>{-r, sg * r});
    o[7] = cmul(o[7], C2<T>{-c1, sg * s1});
    dft8<S>(e);
    dft8<S>(o);
#pragma unroll
    for (int m = 0; m < 8; ++m) { v[2 * m] = e[m]; v[2 * m + 1] = o[m]; }
}

template <int R, int S, typename T, bool FUSE = true>
__device__ __forceinline__ void dftR(C2<T>* v) {
    if constexpr (R == 8) dft8<S, T, FUSE>(v);
    else dft16<S>(v);
}

// Per-lane twiddles of one transform size, loaded once per kernel.
//   t1[k1] = pass-1 twiddle for lane (b,c): W_M^{(8b+c) k1} times the lane part w^{8b+c} of the
//            twist, so the forward twist reduces to the register constants w^{64a}
//   t2[k2] = pass-2 twiddle for lane (c,k1): W_64^{c k2}
template <typename T, int R1>
struct LaneTw {
    C2<T> t1[R1];
    C2<T> t2[8];
};

// The same twiddles with the pass-2 set read from LDS at each use instead of held in registers:
// t2 points at this lane's eight W_64^{c k2} (c = lane mod 8), shared by every lane with that c.
template <typename T, int R1>
struct LaneTwS {
    C2<T> t1[R1];
    const C2<T>* t2;
};

template <typename T, int R1>
__device__ __forceinline__ void load_lane_tw(LaneTw<T, R1>& tw, const C2<T>* __restrict__ g_t1,
                                             const C2<T>* __restrict__ g_t2, int lane) {
#pragma unroll
    for (int r = 0; r < R1; ++r) tw.t1[r] = g_t1[lane * R1 + r];
#pragma unroll
    for (int r = 0; r < 8; ++r) tw.t2[r] = g_t2[lane * 8 + r];
}

// Exchange fence for the transforms.  Every transform is owned by one wavefront and its
// workgroup is that wavefront alone; LDS operations of a wavefront execute in program order, so
// ordering the compiler (a wavefront-scope fence emits no instruction) is all that is needed --
// no s_barrier and no lgkmcnt(0) drain between the exchange write and read.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS exchange addresses.  Each access of a wavefront (one register index, 64 lanes) should touch
// distinct banks within every group of lanes the LDS serves per cycle (16 or 32 lanes for 8-byte
// elements, 8 or 16 for 16-byte elements, with 32 or 64 banks), and the address must stay a
// per-lane base plus a per-register constant so it folds into the instruction offset.
//
// exchange 1: element (k1, b, c) -- written by lane 8b+c (register k1), read by lane c+8(k1 mod 8)
// (register b).  Rows of 64 are padded to 72, so the 8-lane runs of consecutive k1 start 8 slots
// apart modulo any bank width.
constexpr int X1_ROW = 72;
__device__ __forceinline__ int x1_addr(int k1, int b, int c) { return k1 * X1_ROW + 8 * b + c; }
// exchange 2: element (k1 = k1lo + 8 s, k2, c) -- written by lane c+8k1lo (register k2), read by
// lane k1lo+8k2 (register c).  The c ^ k1lo swizzle separates the k1lo of one read, and rows
// padded to 72 separate its k2 (a 64-element stride would put every k2 on the same banks).
__device__ __forceinline__ int x2_addr(int s, int k2, int k1lo, int c) {
    return (s * 8 + k2) * X1_ROW + 8 * k1lo + (c ^ k1lo);
}
// LDS elements an exchange buffer needs for an M-point transform
template <int M>
__host__ __device__ constexpr int xbuf_elems() { return (M / 64) * X1_ROW > M ? (M / 64) * X1_ROW : M; }

// One LDS exchange: register r of every lane is stored at address wr(r), then register r is
// loaded from address rd(r).  HALF exchanges the real and imaginary parts one after the other
// through a buffer of scalars, halving the LDS footprint for the same bytes moved.
template <int R, bool HALF, typename T, typename WR, typename RD>
__device__ __forceinline__ void exchange(C2<T>* v, void* buf, WR wr, RD rd) {
    if constexpr (!HALF) {
        C2<T>* b = (C2<T>*)buf;
#pragma unroll
        for (int r = 0; r < R; ++r) b[wr(r)] = v[r];
        wave_sync();
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = b[rd(r)];
        wave_sync();
    } else {
        T* b = (T*)buf;
#pragma unroll
        for (int r = 0; r < R; ++r) b[wr(r)] = v[r].x;
        wave_sync();
#pragma unroll
        for (int r = 0; r < R; ++r) v[r].x = b[rd(r)];
        wave_sync();
#pragma unroll
        for (int r = 0; r < R; ++r) b[wr(r)] = v[r].y;
        wave_sync();
#pragma unroll
        for (int r = 0; r < R; ++r) v[r].y = b[rd(r)];
        wave_sync();
    }
}

// Forward transform of one wavefront-owned polynomial.  In: v[a] = z_{64a + lane} (already
// multiplied by the a-dependent twist constant when the twist is folded into tw.t1).  Out: v in
// lane order.  buf: xbuf_elems<M>() elements of C2<T> (of T when HALF) private to the wavefront.
template <int R1, typename T, bool HALF = false, bool FUSE = true, typename TW>
__device__ __forceinline__ void fft_fwd(C2<T>* v, const TW& tw, void* vbuf, int lane) {
    constexpr int G = R1 / 8;
    dftR<R1, +1, T, FUSE>(v);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) v[k1] = cmul(v[k1], tw.t1[k1]);
    const int c = lane & 7, k1lo = lane >> 3;
    if constexpr (!HALF) {
        C2<T>* buf = (C2<T>*)vbuf;
        {
            const int b1 = lane >> 3, c1 = lane & 7;
#pragma unroll
            for (int k1 = 0; k1 < R1; ++k1) buf[x1_addr(k1, b1, c1)] = v[k1];
        }
        wave_sync();
#pragma unroll
        for (int s = 0; s < G; ++s)
#pragma unroll
            for (int b = 0; b < 8; ++b) v[s * 8 + b] = buf[x1_addr(k1lo + 8 * s, b, c)];
        wave_sync();
#pragma unroll
        for (int s = 0; s < G; ++s) {
            dft8<+1, T, FUSE>(v + s * 8);
#pragma unroll
            for (int k2 = 1; k2 < 8; ++k2) v[s * 8 + k2] = cmul(v[s * 8 + k2], tw.t2[k2]);
#pragma unroll
            for (int k2 = 0; k2 < 8; ++k2) buf[x2_addr(s, k2, k1lo, c)] = v[s * 8 + k2];
        }
        wave_sync();
        {
            const int k1lo3 = lane & 7, k2 = lane >> 3;
#pragma unroll
            for (int s = 0; s < G; ++s)
#pragma unroll
                for (int cc = 0; cc < 8; ++cc) v[s * 8 + cc] = buf[x2_addr(s, k2, k1lo3, cc)];
        }
        wave_sync();
    } else {
        // lane = 8 b1 + c1 = 8 k1lo + c, so both sides of exchange 1 index with (k1lo, c)
        exchange<R1, true>(v, vbuf, [&](int k1) { return x1_addr(k1, k1lo, c); },
                           [&](int r) { return x1_addr(k1lo + 8 * (r >> 3), r & 7, c); });
#pragma unroll
        for (int s = 0; s < G; ++s) {
            dft8<+1, T, FUSE>(v + s * 8);
#pragma unroll
            for (int k2 = 1; k2 < 8; ++k2) v[s * 8 + k2] = cmul(v[s * 8 + k2], tw.t2[k2]);
        }
        const int k1lo3 = lane & 7, k2r = lane >> 3;
        exchange<R1, true>(v, vbuf, [&](int r) { return x2_addr(r >> 3, r & 7, k1lo, c); },
                           [&](int r) { return x2_addr(r >> 3, k2r, k1lo3, r & 7); });
    }
#pragma unroll
    for (int s = 0; s < G; ++s) dft8<+1, T, FUSE>(v + s * 8);
}

// Inverse of fft_fwd without the 1/M scale and without the a-dependent untwist constants:
// in: v in lane order; out: v[a] = M * z_{64a+lane} * (conj of the forward a-constant).
template <int R1, typename T, bool HALF = false, bool FUSE = true, typename TW>
__device__ __forceinline__ void fft_inv(C2<T>* v, const TW& tw, void* vbuf, int lane) {
    constexpr int G = R1 / 8;
    // conj(t2) on the inputs of an 8-point inverse DFT; FUSE folds each twiddle into a butterfly
    auto inv_pass2 = [](C2<T>* x, const TW& t) {
        if constexpr (FUSE) {
            dft8_tw<-1, true, true>(x, t.t2);
        } else {
#pragma unroll
            for (int k2 = 1; k2 < 8; ++k2) x[k2] = cmulc(x[k2], t.t2[k2]);
            dft8<-1, T, false>(x);
        }
    };
#pragma unroll
    for (int s = 0; s < G; ++s) dft8<-1, T, FUSE>(v + s * 8);
    const int c = lane & 7, k1lo = lane >> 3;
    if constexpr (!HALF) {
        C2<T>* buf = (C2<T>*)vbuf;
        {
            const int k1lo3 = lane & 7, k2 = lane >> 3;
#pragma unroll
            for (int s = 0; s < G; ++s)
#pragma unroll
                for (int cc = 0; cc < 8; ++cc) buf[x2_addr(s, k2, k1lo3, cc)] = v[s * 8 + cc];
        }
        wave_sync();
#pragma unroll
        for (int s = 0; s < G; ++s)
#pragma unroll
            for (int k2 = 0; k2 < 8; ++k2) v[s * 8 + k2] = buf[x2_addr(s, k2, k1lo, c)];
        wave_sync();
#pragma unroll
        for (int s = 0; s < G; ++s) {
            inv_pass2(v + s * 8, tw);
#pragma unroll
            for (int b = 0; b < 8; ++b) buf[x1_addr(k1lo + 8 * s, b, c)] = v[s * 8 + b];
        }
        wave_sync();
        {
            const int b1 = lane >> 3, c1 = lane & 7;
#pragma unroll
            for (int k1 = 0; k1 < R1; ++k1) v[k1] = buf[x1_addr(k1, b1, c1)];
        }
        wave_sync();
    } else {
        const int k1lo3 = lane & 7, k2r = lane >> 3;
        exchange<R1, true>(v, vbuf, [&](int r) { return x2_addr(r >> 3, k2r, k1lo3, r & 7); },
                           [&](int r) { return x2_addr(r >> 3, r & 7, k1lo, c); });
#pragma unroll
        for (int s = 0; s < G; ++s) inv_pass2(v + s * 8, tw);
        exchange<R1, true>(v, vbuf, [&](int r) { return x1_addr(k1lo + 8 * (r >> 3), r & 7, c); },
                           [&](int k1) { return x1_addr(k1, k1lo, c); });
    }
    if constexpr (FUSE && R1 == 8) {
        dft8_tw<-1, true, false>(v, tw.t1);
    } else {
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) v[k1] = cmulc(v[k1], tw.t1[k1]);
        dftR<R1, -1, T, FUSE>(v);
    }
}

// Twiddles of the decimation-in-time 512-point transform (fft512_dit), per lane L; entry 0 of
// each is 1 and never read.
//   t2[b] = W_64^{b (L >> 3)}        pass-2 inputs, lane L = c + 8 k1
//   t3[c] = W_512^{c L}              pass-3 inputs, lane L = k1 + 8 k2
template <typename T>
struct DitTw {
    C2<T> t2[8];
    C2<T> t3[8];
};

template <typename T>
__device__ __forceinline__ void load_dit_tw(DitTw<T>& tw, const C2<T>* __restrict__ g_t2,
                                            const C2<T>* __restrict__ g_t3, int lane) {
#pragma unroll
    for (int r = 1; r < 8; ++r) {
        tw.t2[r] = g_t2[lane * 8 + r];
        tw.t3[r] = g_t3[lane * 8 + r];
    }
}

// Exchange of lane bit 3 with a register bit for a pair of values (r0: register bit 0, r1: 1):
// each lane sends its partner (lane ^ 8, DPP row_ror:8 within the row of 16) the value the partner
// keeps and takes the partner's in its place.  s3: this lane's bit 3.
#ifndef SWAP_BIT3_BANKMASK
#define SWAP_BIT3_BANKMASK 1
#endif
__device__ __forceinline__ void swap_lane_bit3(uint32_t& r0, uint32_t& r1, bool s3) {
#if SWAP_BIT3_BANKMASK
    // Lanes with bit 3 set are DPP banks 2 and 3 of each row (bank_mask 0xC), the others banks 0
    // and 1 (0x3); a DPP move writes only the lanes of its enabled banks and leaves the others at
    // the old value.  So: r0 of the bit-3 lanes takes the partner's r1, then r1 of the other lanes
    // takes the partner's original r0 (t).  No selects, since neither move needs one.
    (void)s3;
    const uint32_t t = r0;
    r0 = (uint32_t)__builtin_amdgcn_update_dpp((int)r0, (int)r1, 0x128, 0xF, 0xC, false);
    r1 = (uint32_t)__builtin_amdgcn_update_dpp((int)r1, (int)t, 0x128, 0xF, 0x3, false);
#else
    const uint32_t send = s3 ? r0 : r1;
    const uint32_t got = (uint32_t)__builtin_amdgcn_mov_dpp((int)send, 0x128, 0xF, 0xF, true);  // row_ror:8
    r0 = s3 ? got : r0;
    r1 = s3 ? r1 : got;
#endif
}

// 512-point transform with natural order in and out (lane L, register r <-> index L + 64 r), for
// both directions: S = +1 computes F_k = sum_j z_j W_512^{jk}; S = -1 runs the same factorisation
// with every root conjugated, which gives M times the inverse.  Decimation in time over
// j = 64a + 8b + c, k = k1 + 8 k2 + 64 k3, with
//   F_k = sum_c W_8^{c k3} W_512^{c (k1 + 8 k2)} sum_b W_8^{b k2} W_64^{b k1} sum_a W_8^{a k1} z_j :
//   pass 1  lane 8b+c,   registers a -> k1   (TWIST: the inputs are first multiplied by tws[a])
//   pass 2  lane c+8k1,  registers b -> k2,  inputs times W_64^{b k1}
//   pass 3  lane k1+8k2, registers c -> k3,  inputs times W_512^{c (k1+8k2)}
// The exchanges are those of fft_fwd; every twiddle (and the twist) multiplies a butterfly input
// and is fused into it (dft8_tw), so a transform costs 196 real operations per lane against 212
// for fft_fwd with output twiddles.  buf: xbuf_elems<512>() elements private to the wavefront.
template <int S, bool TWIST, typename T>
__device__ __forceinline__ void fft512_dit(C2<T>* v, const DitTw<T>& tw, const C2<T>* tws, C2<T>* buf, int lane) {
    constexpr bool CJ = S < 0;
    if constexpr (TWIST) dft8_tw<S, false, false>(v, tws);
    else dft8<S, T>(v);
    const int lo = lane & 7, hi = lane >> 3;
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1) buf[x1_addr(k1, hi, lo)] = v[k1];
    wave_sync();
#pragma unroll
    for (int b = 0; b < 8; ++b) v[b] = buf[x1_addr(hi, b, lo)];
    wave_sync();
    dft8_tw<S, CJ, true>(v, tw.t2);
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) buf[x2_addr(0, k2, hi, lo)] = v[k2];
    wave_sync();
#pragma unroll
    for (int c = 0; c < 8; ++c) v[c] = buf[x2_addr(0, hi, lo, c)];
    wave_sync();
    dft8_tw<S, CJ, true>(v, tw.t3);
}

// ---- forward transform of gadget digits with its first pass on the matrix cores ----
//
// The forward transform's inputs are gadget digits: integers of at most 11 bits, exact in FP16.  Its
// first pass (fft512_dit: the 8-point DFT over a of z_{64a+L} w^{64a+L}) is, per vector L, the
// fixed complex 8x8 matrix G[k1][a] = W_8^{a k1} w^{64a} applied to the eight inputs, followed by
// the vector's own factor w^L.  As a real 16x16 matrix A (rows 2 k1 + {re, im}, columns
// 2 a + {re, im}) it multiplies 16 vectors at once on the matrix cores.  A is split into two FP16
// pieces, A1 = fp16(A) and A2 = fp16(A - A1): the entries lie in [-1, 1], so |A - A1| <= 2^-12 is a
// normal FP16 number and A1 + A2 is within 2^-24 of A, the FP32 level.  The digits are exact, the
// products of FP16 numbers are exact in FP32, and the sums accumulate in FP32.  The factor
// w^L = w^{8b + c} moves into the pass-2 input twiddles (fwd_t2[b] = W_64^{b k1} w^{8b+c} for lane
// c + 8 k1; b = 0 is no longer 1).  A wavefront's 64 vectors are four groups of 16.
//
// Two instruction forms (FWD_MFMA_K32):
//   1: one v_mfma_f32_16x16x32_f16 per group, the pieces stacked along K (columns 0-15 of the
//      operand are A1, 16-31 are A2, and both halves of B hold the same 16 inputs);
//   0: two v_mfma_f32_16x16x16_f16 per group, one per piece, accumulating into the same output.
// Operand layouts (checked on gfx950), q = l >> 4, n0 = l & 15: K = 16: lane l holds
// A[n0][4q + i] and B[4q + i][n0], i < 4; K = 32: A[n0][8q + i] and B[8q + i][n0], i < 8; both:
// D[4q + i][n0], i < 4.  So lane (n0, q) of group g gets outputs k1 = 2q, 2q + 1 of vector 16g + n0.
#ifndef FWD_MFMA_K32
#define FWD_MFMA_K32 0
#endif
// 1: the matrix-core outputs reach pass 2's layout in registers (one lane-bit exchange), 0: through
// the LDS exchange buffer
#ifndef FWD_MFMA_X1_REGS
#define FWD_MFMA_X1_REGS 1
#endif
typedef _Float16 f16x4_t __attribute__((ext_vector_type(4)));
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

// This lane's entries of the pass-1 matrix pieces.  Table [64 lanes][16] FP16: entries 0-3 piece 1
// and 4-7 piece 2 in the K = 16 layout, entries 8-15 the stacked K = 32 layout.
struct FwdMfmaA {
#if FWD_MFMA_K32
    f16x8_t a;
#else
    f16x4_t a[2];
#endif
};

__device__ __forceinline__ void load_fwd_mfma_a(FwdMfmaA& m, const f16x8_t* __restrict__ tab, int lane) {
#if FWD_MFMA_K32
    m.a = tab[lane * 2 + 1];
#else
    const f16x8_t t = tab[lane * 2];
    m.a[0] = f16x4_t{t[0], t[1], t[2], t[3]};
    m.a[1] = f16x4_t{t[4], t[5], t[6], t[7]};
#endif
}

// two FP32 integers as an FP16 pair (exact below 2^11), x in the low half
__device__ __forceinline__ uint32_t pack_f16x2(float x, float y) {
    return __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(x, y));
}

// Gadget digits straight to FP16 pairs, from y = x + offset (no xor): each Bg-bit field of y is
// f = digit + 2^(Bg-1), unsigned.  The field goes into the mantissa of an FP16 number whose
// exponent makes one unit of f one unit of value, and one packed FP16 add subtracts the exponent's
// base and 2^(Bg-1), exactly (every value is an integer below 2^11).  Replaces a signed bit-field
// extract and an int-to-float conversion per digit and a packing conversion per pair.
typedef _Float16 f16x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t f16x2_add(uint32_t w, _Float16 c) {
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(f16x2_t, w) + f16x2_t{c, c});
}
// Bg = 8: level lvl is byte 3 - lvl of a (low half) and b (high half); with 0x64 (1024 + f) above
// each byte, minus 1024 + 128.
__device__ __forceinline__ uint32_t digits8_f16x2(uint32_t a, uint32_t b, int lvl) {
    const uint32_t k = 3u - (uint32_t)lvl;
    const uint32_t w = __builtin_amdgcn_perm(b, a, k | (0x0Cu << 8) | ((4u + k) << 16) | (0x0Cu << 24));
    return f16x2_add(w | 0x64006400u, (_Float16)-1152.0f);
}
// Bg = 7: the top two bytes of a and b (p0 = byte-permuted (a >> 16, b >> 16)) hold the fields of
// levels 0 (bits 9-15 of each half) and 1 (bits 2-8), bytes 1-2 (p1) the field of level 2 (bits
// 3-9).  Each is moved to mantissa bits 3-9 of exponent 0x58 (128 + f, steps of 1/8) by a 32-bit
// shift, whose bits crossing between the halves the mask drops, then 128 + 64 is subtracted.
__device__ __forceinline__ void digits7_prep(uint32_t a, uint32_t b, uint32_t& p0, uint32_t& p1) {
    p0 = __builtin_amdgcn_perm(b, a, 0x07060302u);
    p1 = __builtin_amdgcn_perm(b, a, 0x06050201u);
}
__device__ __forceinline__ uint32_t digits7_f16x2(uint32_t p0, uint32_t p1, int lvl) {
    const uint32_t h = lvl == 0 ? p0 >> 6 : lvl == 1 ? p0 << 1 : p1;
    return f16x2_add((h & 0x03F803F8u) | 0x58005800u, (_Float16)-192.0f);
}
// The same for one level only, with the one permute that level needs.
__device__ __forceinline__ uint32_t digits7_f16x2_one(uint32_t a, uint32_t b, int lvl) {
    const uint32_t p = __builtin_amdgcn_perm(b, a, lvl < 2 ? 0x07060302u : 0x06050201u);
    const uint32_t h = lvl == 0 ? p >> 6 : lvl == 1 ? p << 1 : p;
    return f16x2_add((h & 0x03F803F8u) | 0x58005800u, (_Float16)-192.0f);
}

// Lane bits 4 and 5 trade places with register bits 1 and 2 of dw[8] (v_permlane16_swap of
// registers (r, r + 2), then v_permlane32_swap of (r, r + 4)).
__device__ __forceinline__ void swap_lane45_reg12(uint32_t* dw) {
#pragma unroll
    for (int r : {0, 1, 4, 5}) {
        const auto t = __builtin_amdgcn_permlane16_swap(dw[r], dw[r + 2], false, false);
        dw[r] = t[0];
        dw[r + 2] = t[1];
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const auto t = __builtin_amdgcn_permlane32_swap(dw[r], dw[r + 4], false, false);
        dw[r] = t[0];
        dw[r + 4] = t[1];
    }
}

// MFMA accumulators d[g] (lane (n0, q) holds (k1 = 2q + t) of vector 16g + n0 in d[g][2t],
// d[g][2t + 1]) to pass 2's layout (lane c + 8 k1 holds every b of its k1, v[b]): one lane-bit
// exchange (lane bit 3 with the register bit t, swap_lane_bit3).
__device__ __forceinline__ void mfma_out_to_pass2(C2<float>* v, const f32x4_t* d, int lane) {
    const bool s3 = (lane & 8) != 0;
#pragma unroll
    for (int g = 0; g < 4; ++g)
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            // (__builtin_bit_cast of an element of an ext_vector_type array reads element 0; the
            // element is taken by value here)
            uint32_t r0 = __float_as_uint(d[g][i]), r1 = __float_as_uint(d[g][2 + i]);
            swap_lane_bit3(r0, r1, s3);
            (i ? v[2 * g].y : v[2 * g].x) = __uint_as_float(r0);
            (i ? v[2 * g + 1].y : v[2 * g + 1].x) = __uint_as_float(r1);
        }
}

// Forward 512-point transform of digit inputs.  In: dw[a] = FP16 pair (Re, Im) of input a of this
// lane's vector (the digits of coefficients L + 64a and L + 64a + 512), i.e. fft512_dit's inputs
// before the twist; dw is overwritten.  Out: v in natural order, as fft512_dit<+1, true> with the
// twist w^{64a+L}.  fwd_t2: the pass-2 table with the lane twist folded in; tw.t3 as for fft512_dit.
// The operands move between lanes with v_permlane16_swap (rows 1 and 3 of its first operand
// exchanged with rows 0 and 2 of its second; rows of 16 lanes) and v_permlane32_swap (the upper
// half of its first operand with the lower half of its second).  The outputs reach pass 2's layout
// by one lane-bit exchange in registers (FWD_MFMA_X1_REGS) instead of exchange 1 through buf, which
// is then used only by exchange 2.
// IN_B: dw is already in the operand layout (lane (n0, q), register 2g + t: input 2q + t of vector
// 16g + n0), as the bootstrap kernel reads it from the accumulator.
template <bool IN_B = false>
__device__ __forceinline__ void fft512_fwd_mfma(C2<float>* v, uint32_t* dw, const FwdMfmaA& am,
                                                const C2<float>* fwd_t2, const DitTw<float>& tw, C2<float>* buf,
                                                int lane) {
    static_assert(!(IN_B && FWD_MFMA_K32), "the operand-layout input is for the K = 16 form");
    f32x4_t d[4];
#if FWD_MFMA_K32
    // group g, lane (n0, q): the four inputs 4 (q & 1) + i of vector 16g + n0, from row g.  The swap
    // of (dw[i], dw[4+i]) puts row r's dw[i] and dw[4+i] in rows (0, 1) of its first result (r even)
    // or second (r odd), and of rows (2, 3) for r + 2; a self-swap of 32 then copies each half to
    // both halves.
    uint32_t bg[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const auto s16 = __builtin_amdgcn_permlane16_swap(dw[i], dw[4 + i], false, false);
        const auto e = __builtin_amdgcn_permlane32_swap(s16[0], s16[0], false, false);
        const auto o = __builtin_amdgcn_permlane32_swap(s16[1], s16[1], false, false);
        bg[0][i] = e[0];
        bg[2][i] = e[1];
        bg[1][i] = o[0];
        bg[3][i] = o[1];
    }
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const f16x8_t b = __builtin_bit_cast(f16x8_t, (uint4){bg[g][0], bg[g][1], bg[g][2], bg[g][3]});
        d[g] = __builtin_amdgcn_mfma_f32_16x16x32_f16(am.a, b, f32x4_t{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
    }
#else
    // lane bits 4, 5 <-> register bits 1, 2 of dw: lane (n0, q) then holds inputs 2q, 2q + 1 of
    // vector 16g + n0 in dw[2g], dw[2g + 1]
    if constexpr (!IN_B) swap_lane45_reg12(dw);
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const f16x4_t b = __builtin_bit_cast(f16x4_t, (uint2){dw[2 * g], dw[2 * g + 1]});
        d[g] = __builtin_amdgcn_mfma_f32_16x16x16f16(am.a[1], b, f32x4_t{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
        d[g] = __builtin_amdgcn_mfma_f32_16x16x16f16(am.a[0], b, d[g], 0, 0, 0);
    }
#endif
    const int lo = lane & 7, hi = lane >> 3;
#if FWD_MFMA_X1_REGS
    // Lane (n0, q) holds (k1 = 2q + t, b = 2g + (n0 >> 3)) in d[g][2t], d[g][2t + 1]; pass 2 wants lane
    // c + 8 k1 to hold every b of its k1.  Only lane bit 3 (b's low bit here, k1's low bit there)
    // and the register bit t trade places.
    mfma_out_to_pass2(v, d, lane);
#else
    const int q = lane >> 4, n0 = lane & 15;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        // element (k1, vector 8b + c) of exchange 1 sits at x1_addr(k1, b, c) = 72 k1 + 8b + c
        buf[(2 * q) * X1_ROW + 16 * g + n0] = C2<float>{d[g][0], d[g][1]};
        buf[(2 * q + 1) * X1_ROW + 16 * g + n0] = C2<float>{d[g][2], d[g][3]};
    }
    wave_sync();
#pragma unroll
    for (int b = 0; b < 8; ++b) v[b] = buf[x1_addr(hi, b, lo)];
    wave_sync();
#endif
    dft8_tw<+1, false, false>(v, fwd_t2);
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) buf[x2_addr(0, k2, hi, lo)] = v[k2];
    wave_sync();
#pragma unroll
    for (int c = 0; c < 8; ++c) v[c] = buf[x2_addr(0, hi, lo, c)];
    wave_sync();
    dft8_tw<+1, false, true>(v, tw.t3);
}

// N = 2048, split path: the 512-point forward transform of wavefront v's half of the spectrum,
// y_v[j'] = w^{(1+4v) j'} (P_{j'} + c_v P_{j'+512}), P_j = p_j + i p_{j+1024}, c_v = (-1)^v e^{i pi/4}
// (the decimation-in-frequency fft_fwd<8> with its pass-1 twiddles w^{L(1+4v+8k1)}).  The
// combination P0 + c_v P1, the register twist w^{64a(1+4v)} and the first 8-point pass form, per
// vector L, a fixed complex 8x16 matrix on the 16 digit pairs (P0_a, P1_a); as a real 16x32 matrix
// (rows 2 k1 + {re, im}, columns 4a + {P0 re, P0 im, P1 re, P1 im}) split into two FP16 pieces it
// runs as two v_mfma_f32_16x16x32_f16 per group of 16 vectors.  Operand layouts (checked on
// gfx950), q = l >> 4, n0 = l & 15: lane l holds A[n0][8q + i] and B[8q + i][n0], i < 8, and
// D[4q + i][n0], i < 4.  In: p0[a], p1[a] = FP16 pairs of P0_a, P1_a of this lane's vector
// (overwritten).  a2: this lane's entries of the two pieces.  t1p: the pass-1 twiddles in pass 2's
// layout (lane c + 8 k1, register b: w^{L(1+4v+8k1)}, L = 8b + c), applied on pass 2's inputs.
// Out: v in lane order, as fft_fwd<8> of y_v.  IN_B: the inputs are already in the operand layout
// (lane (n0, q), register 2g + t: input 2q + t of vector 16g + n0), as the bootstrap kernel reads
// them from the accumulator; otherwise lane L holds input a of vector L in register a.
template <bool IN_B = false>
__device__ __forceinline__ void fft512_fwd2w_mfma(C2<float>* v, uint32_t* p0, uint32_t* p1, const f16x8_t* a2,
                                                  const C2<float>* t1p, const C2<float>* t2, C2<float>* buf,
                                                  int lane) {
    // lane (n0, q) of group g needs inputs 2q, 2q + 1 of vector 16g + n0
    if constexpr (!IN_B) {
        swap_lane45_reg12(p0);
        swap_lane45_reg12(p1);
    }
    f32x4_t d[4];
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const f16x8_t b = __builtin_bit_cast(f16x8_t, (uint4){p0[2 * g], p1[2 * g], p0[2 * g + 1], p1[2 * g + 1]});
        d[g] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a2[1], b, f32x4_t{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
        d[g] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a2[0], b, d[g], 0, 0, 0);
    }
    mfma_out_to_pass2(v, d, lane);
    const int c = lane & 7, k1lo = lane >> 3;
    dft8_tw<+1, false, false>(v, t1p);
#pragma unroll
    for (int k2 = 1; k2 < 8; ++k2) v[k2] = cmul(v[k2], t2[k2]);
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) buf[x2_addr(0, k2, k1lo, c)] = v[k2];
    wave_sync();
    {
        const int k1lo3 = lane & 7, k2 = lane >> 3;
#pragma unroll
        for (int cc = 0; cc < 8; ++cc) v[cc] = buf[x2_addr(0, k2, k1lo3, cc)];
    }
    wave_sync();
    dft8<+1, float>(v);
}

// The inverse of fft512_fwd2w_mfma's transform (fft_inv<8> of the split N = 2048 path) with the
// conjugate pass-1 twiddles taken from the same pass-2-layout table t1p, applied before exchange 1
// instead of on the last pass's inputs.
__device__ __forceinline__ void fft512_inv2w_t1p(C2<float>* v, const C2<float>* t1p, const C2<float>* t2,
                                                 C2<float>* buf, int lane) {
    dft8<-1, float>(v);
    const int c = lane & 7, k1lo = lane >> 3;
    {
        const int k1lo3 = lane & 7, k2 = lane >> 3;
#pragma unroll
        for (int cc = 0; cc < 8; ++cc) buf[x2_addr(0, k2, k1lo3, cc)] = v[cc];
    }
    wave_sync();
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) v[k2] = buf[x2_addr(0, k2, k1lo, c)];
    wave_sync();
    dft8_tw<-1, true, true>(v, t2);
#pragma unroll
    for (int b = 0; b < 8; ++b) v[b] = cmulc(v[b], t1p[b]);
#pragma unroll
    for (int b = 0; b < 8; ++b) buf[x1_addr(k1lo, b, c)] = v[b];
    wave_sync();
    {
        const int b1 = lane >> 3, c1 = lane & 7;
#pragma unroll
        for (int k1 = 0; k1 < 8; ++k1) v[k1] = buf[x1_addr(k1, b1, c1)];
    }
    wave_sync();
    dft8<-1, float>(v);
}

// ---- integer stages ----

// coefficient j of (X^a - 1) * p, a in [0, 2N): e = (j - a) mod 2N; (X^a p)_j = p_e or -p_{e-N}
// (kernel/src/tfhe_blindRotate_FFT/tfhe_MuxRotate_FFT.cpp:5-24)
template <int N>
__device__ __forceinline__ uint32_t rotdiff(const int32_t* p, int j, int a) {
    int e = (j - a) & (2 * N - 1);
    uint32_t s = (uint32_t)p[e & (N - 1)];
    s = (e < N) ? s : (0u - s);
    return s - (uint32_t)p[j];
}

// The same coefficient plus `off` (the gadget offset of the decomposition that follows), written
// for the instruction count of the blind-rotation loop.  Coefficient j = lane + 64 r takes the byte
// offset e4 = base4 + 256 r, base4 = 4 (lane - a) (rot_base4, once per step): the address is
// e4 mod 4N, and m = -1 when (j - a) mod 2N >= N (bit log2(4N) of e4, one signed bit-field
// extract), so that +-p_e = (p_e ^ m) - m.
__device__ __forceinline__ uint32_t rot_base4(int lane, int a) { return (uint32_t)(lane - a) << 2; }

template <int N>
__device__ __forceinline__ uint32_t rotdiff_e4(const int32_t* p, uint32_t e4, int j, uint32_t off) {
    constexpr int SH = __builtin_ctz(4 * N);
    const uint32_t s = *(const uint32_t*)((const char*)p + (e4 & (4u * N - 4u)));
    const uint32_t m = (uint32_t)((int32_t)(e4 << (31 - SH)) >> 31);
    return (s ^ m) + ((off - m) - (uint32_t)p[j]);
}

// signed gadget digit of level lvl (tGswTorus32PolynomialDecompH.cpp:29-44) of u = x + offset
template <int BG>
__device__ __forceinline__ int digit(uint32_t u_with_offset, int lvl) {
    return (int)((u_with_offset >> (32 - (lvl + 1) * BG)) & ((1u << BG) - 1)) - (1 << (BG - 1));
}

// The same digit from ux = (x + offset) ^ offset.  The offset is 2^(BG-1) in every BG-bit field,
// so the xor flips each field's top bit, and field - 2^(BG-1) is then the field read as a signed
// BG-bit number: one signed bit-field extract instead of an extract and a subtract.
template <int BG>
__device__ __forceinline__ int digit_sx(uint32_t ux, int lvl) {
    return ((int32_t)(ux << (lvl * BG))) >> (32 - BG);
}

template <int L, int BG>
__host__ __device__ constexpr uint32_t decomp_offset() {
    uint32_t off = 0;
    for (int p = 0; p < L; ++p) off += (1u << (BG - 1)) * (1u << (32 - (p + 1) * BG));
    return off;
}

// round an integer-valued double (|x| < 2^51) to nearest and reduce mod 2^32:
// adding 1.5 * 2^52 puts the unit in the last mantissa bit, so the low word is the residue
__device__ __forceinline__ uint32_t round_mod32(double x) {
    double y = x + 6755399441055744.0;
    return (uint32_t)__double_as_longlong(y);
}

}  // namespace tfhe
