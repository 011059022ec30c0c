// gpp_ozaki_dev.hpp -- device code of the Ozaki Gram product (the split of
// both operands into digit planes and the layout-3 tile on 32x32 MFMAs),
// shared by the single-instance launches of gpp_ozaki.hip and the batched
// launches of gpp_batch.hip.  The scheme, its exactness bounds and the MFMA
// lane layouts are described at the top of gpp_ozaki.hip.
#pragma once

#include "gpp_internal.hpp"

namespace gpp {
namespace oz {

typedef _Float16 v8h __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef int v4i __attribute__((ext_vector_type(4)));
typedef signed char v16c __attribute__((ext_vector_type(16)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int kOzTile = 64;  // padding of both operands' columns

// Digit-type traits: FP16 digits on f32_16x16x32_f16 with FP32 sums over
// chunks of 256 bands (exact for levels 0-2); INT8 digits on i32_16x16x64_i8
// with exact INT32 sums.
template <typename D> struct OzDigit;
template <> struct OzDigit<_Float16> {
    typedef v8h frag;
    typedef v4f acc;
    static constexpr int KS = 32;       // bands per MFMA k-step = per LDS stage (64 bytes)
    // Bands per FP32 accumulation chunk (OzAcc: the levels exact over it).
    static constexpr int CHUNK = 256;
    static constexpr bool ONE_CHUNK = false;
    static __device__ __forceinline__ acc mfma(frag a, frag b, acc c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
    // 32 x 32 tiles (layout 3): f32_32x32x16_f16, 16 bands per k-step.
    typedef v16f acc32;
    static constexpr int KS32 = 16;
    static __device__ __forceinline__ acc32 mfma32(frag a, frag b, acc32 c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
    // Sign flip of the eight halves (exact negation), one XOR per dword.
    static __device__ __forceinline__ frag neg(frag x)
    {
        const v4i m = {(int)0x80008000, (int)0x80008000, (int)0x80008000, (int)0x80008000};
        return __builtin_bit_cast(frag, __builtin_bit_cast(v4i, x) ^ m);
    }
};
template <> struct OzDigit<int8_t> {
    typedef v4i frag;
    typedef v4i acc;
    static constexpr int KS = 64;
    // INT32 sums are exact over 8192 bands (a level takes at most
    // 2 (2*64*127 + 4*127^2) = 161544 per band at 6 slices; x 8192 < 2^31):
    // band ranges up to that are one chunk, folded to FP64 once at the end;
    // longer ones run the tile instantiated for long ranges (LONG), which
    // flushes to FP64 every 8192 bands like the FP16 chunks.
    static constexpr int CHUNK = 8192;
    static constexpr bool ONE_CHUNK = true;
    static __device__ __forceinline__ acc mfma(frag a, frag b, acc c)
    {
        return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
    }
    typedef v16i acc32;
    static constexpr int KS32 = 32;
    static __device__ __forceinline__ acc32 mfma32(frag a, frag b, acc32 c)
    {
        return __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c, 0, 0, 0);
    }
    // Byte-wise two's complement negation, four bytes per dword without
    // carries between bytes: -x = (~x) + 1 per byte.
    static __device__ __forceinline__ frag neg(frag x)
    {
        frag r;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const unsigned t = ~(unsigned)x[i];
            r[i] = (int)(((t & 0x7f7f7f7fu) + 0x01010101u) ^ (t & 0x80808080u));
        }
        return r;
    }
};

// Digits and accumulators of S slices of digit type D.
//
// Digits: a real part x of a column with exponent e is
//   x = 2^(e-6) (d_0 + d_1 / B + d_2 / B^2 + ...),
// d_0 in [-64, 64] (x 2^(6-e) rounded), d_p for p >= 1 in [-B/2, B/2] (B times
// the previous remainder, at most 0.5 in magnitude, rounded).  B = 254 (about
// 8 bits per digit after the first: 39 bits at 5 slices, 47 at 6; exact in
// int8 and FP16), except FP16 at 6 slices, which keeps B = 128 (42 bits): its
// wider digits would need a fourteenth accumulator, beyond the registers.
//
// Accumulators: one per level t = p + q of the digit pairs (p, q), folded
// with weight B^-t.  Per band a level takes (t+1) pairs x 2 products.  INT8
// sums are exact in INT32 over a chunk.  FP16 sums, in FP32, are exact while
// below 2^24 -- over 256 bands: with B = 128 every level (6 * 2 * 64^2 * 256
// < 2^24); with B = 254 levels 0-2 (level 2: 2 (2*64*127 + 127^2) * 256 =
// 16581120) and level 3 split over two accumulators (pairs p <= 1 and p >= 2,
// each 2 (64*127 + 127^2) * 256 = 12419584), while levels 4 and up round,
// within 2^-40 of the chunk's largest possible sum given their weights
// (254^-4 and below, half an ulp of 2^26 per product at most).
template <typename D, int S> struct OzAcc {
    static constexpr bool WIDE = !(sizeof(D) == 2 && S == 6);  // B = 254
    static constexpr double BASE = WIDE ? 254.0 : 128.0;
    static constexpr double INV_BASE = 1.0 / BASE;
    static constexpr bool SPLIT = sizeof(D) == 2 && WIDE && S >= 4;
    static constexpr int N = SPLIT ? S + 1 : S;  // accumulators
    static constexpr __host__ __device__ int idx(int p, int q)
    {
        return !SPLIT || p + q < 3 ? p + q : p + q == 3 ? (p < 2 ? 3 : 4) : p + q + 1;
    }
    // accumulator of level t (its first half where split)
    static constexpr __host__ __device__ int first(int t) { return !SPLIT || t <= 3 ? t : t + 1; }
};

// 16 bytes per lane from global memory straight into LDS: the wave writes
// 1 KB at lds_wave (wave-uniform) + 16 * lane.
__device__ __forceinline__ void load_lds16(const void *src, void *lds_wave)
{
    __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void *)lds_wave, 16, 0, 0);
}

// Workgroup barrier without the memory fence of __syncthreads() (whose
// release would wait for every outstanding load, the prefetched stages
// included); the caller orders its memory with explicit s_waitcnt.  The empty
// asm statements keep the compiler from moving memory accesses across it.
__device__ __forceinline__ void barrier_no_fence()
{
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

__host__ __device__ inline int round_up(int x, int m) { return (x + m - 1) / m * m; }
__host__ inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Column exponent e with 2^e > max over rows [k_begin, k_end) of |re|, |im| of
// column c of a row-major complex matrix (row stride ld).  Zero columns get 0.
// One 256-thread block: columns [64 cb, 64 cb + 64) x 4 row groups (rows
// strided by 4), combined in `part` (LDS).
__device__ __forceinline__ void colexp_block(const cplx *__restrict__ src, int ld, int k_begin, int k_end, int ncols,
                                             int ncols_pad, int *__restrict__ exps, int cb, double (*part)[64])
{
    const int cl = threadIdx.x % 64, rg = threadIdx.x / 64;
    const int c = cb * 64 + cl;
    double m0 = 0.0, m1 = 0.0;
    if (c < ncols) {
        int k = k_begin + rg;
        for (; k + 4 < k_end; k += 8) {
            const cplx v = src[(size_t)k * ld + c], w = src[(size_t)(k + 4) * ld + c];
            m0 = fmax(m0, fmax(fabs(v.re), fabs(v.im)));
            m1 = fmax(m1, fmax(fabs(w.re), fabs(w.im)));
        }
        if (k < k_end) {
            const cplx v = src[(size_t)k * ld + c];
            m0 = fmax(m0, fmax(fabs(v.re), fabs(v.im)));
        }
    }
    part[rg][cl] = fmax(m0, m1);
    __syncthreads();
    if (rg == 0 && c < ncols_pad) {
        const double m = fmax(fmax(part[0][cl], part[1][cl]), fmax(part[2][cl], part[3][cl]));
        int e = 0;
        if (m > 0.0) (void)frexp(m, &e);  // m = f 2^e, f in [0.5, 1): 2^e > m
        exps[c] = e;
    }
}

// Digits of column c, bands [k_begin + KP*kq, +KP) (KP = 16 bytes of digits),
// into the 16-byte piece out + (((p*2 + part) * (kpad/KP) + kq) * ncols_pad + c) * KP:
// consecutive columns' pieces are contiguous, so both the column-wise reads of
// src and the writes are coalesced.  One thread per (column c, piece kq).
template <int S, typename D>
__device__ __forceinline__ void split_piece(const cplx *__restrict__ src, int ld, int k_begin, int K, int ncols,
                                            int ncols_pad, int kpad, const int *__restrict__ exps,
                                            D *__restrict__ out, int c, int kq)
{
    constexpr int KP = 16 / (int)sizeof(D);
    const int e = exps[c];
    D dre[S][KP], dim[S][KP];
#pragma unroll
    for (int kk = 0; kk < KP; ++kk) {
        const int k = kq * KP + kk;
        cplx v = czero();
        if (c < ncols && k < K) {  // read once here: kept out of the caches, which the digits use
            const cplx *q = src + (size_t)(k_begin + k) * ld + c;
            v = cplx{__builtin_nontemporal_load(&q->re), __builtin_nontemporal_load(&q->im)};
        }
        double yr = ldexp(v.re, 6 - e), yi = ldexp(v.im, 6 - e);  // |y| < 64, exact
#pragma unroll
        for (int p = 0; p < S; ++p) {
            const double dr = rint(yr), di = rint(yi);
            dre[p][kk] = (D)dr;
            dim[p][kk] = (D)di;
            // |yr - dr| <= 0.5, so the next digit is in [-B/2, B/2]; with
            // B = 254 the product's rounding (2^-53 of the remainder) is far
            // below the split's truncation (with B = 128 it is exact)
            yr = (yr - dr) * OzAcc<D, S>::BASE;
            yi = (yi - di) * OzAcc<D, S>::BASE;
        }
    }
    const size_t npieces = (size_t)(kpad / KP);
#pragma unroll
    for (int p = 0; p < S; ++p) {
        __builtin_memcpy(out + (((size_t)(p * 2 + 0) * npieces + kq) * ncols_pad + c) * KP, dre[p], 16);
        __builtin_memcpy(out + (((size_t)(p * 2 + 1) * npieces + kq) * ncols_pad + c) * KP, dim[p], 16);
    }
}

// Layout 3: Mu tiles of 64 G' columns x 64 G-vectors, four waves of 32 x 32
// on 32x32 MFMAs (one wave per SIMD, so 512 registers each), one k-step of
// bands per stage, NBUF LDS stage buffers (as many as fit, at most four)
// filled straight from memory NBUF - 1 stages ahead, one barrier per stage.
//
// LDS stage buffer: per side (G' rows, then G-vector rows) S planes x 64 rows
// of 64 bytes, each row four 16-byte units (u = 2 part + half: the real part's
// first and second half of the step's bands, then the imaginary part's).
// Unit u of row r sits in slot u ^ ((r / 4) % 4), which makes the fragment
// reads (lane l: row l % 32, half l / 32) free of LDS bank conflicts on
// gfx950 (measured; an XOR with r % 4, or 128-byte lines of two rows, is not).
constexpr int kOz3Tile = 64;
constexpr int kLdsBytes = 160 * 1024;  // LDS per workgroup on gfx950
constexpr int oz3_buffers(int S) { return kLdsBytes / (2 * S * 64 * 64) < 4 ? kLdsBytes / (2 * S * 64 * 64) : 4; }

// s_waitcnt vmcnt(N) (N < 64), leaving the other counters alone.
template <int N> __device__ __forceinline__ void wait_vmcnt()
{
    static_assert(N >= 0 && N < 64, "vmcnt has six bits");
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// The 64 x 64 tile (j0, i0) of Mu (256 threads; lds: the block's NBUF stage
// buffers of Oz3<S>::STAGE bytes).  Returns with every load of the tile waited
// for; a caller that runs further tiles in the same block puts a barrier
// between them.
template <int S> struct Oz3 {
    static constexpr int PLANE = 64 * 64;  // bytes per digit plane in a stage buffer
    static constexpr int SIDE = S * PLANE;
    static constexpr int STAGE = 2 * SIDE;
};
template <int S, typename D, int NBUF, bool LONG = false>
__device__ __forceinline__ void gram_tile32(const D *__restrict__ bd, const D *__restrict__ ad,
                                            const int *__restrict__ eb, const int *__restrict__ ea, int ngpown,
                                            int ncouls, int ngp_pad, int nc_pad, int kpad, int j0, int i0,
                                            cplx *__restrict__ mout, unsigned char (*lds)[Oz3<S>::STAGE])
{
    typedef OzDigit<D> T;
    typedef typename T::frag frag;
    typedef typename T::acc32 acc_t;
    constexpr int KS = T::KS32;                 // bands per stage (32 bytes per part and row)
    constexpr int KP = 16 / (int)sizeof(D);     // bands per 16-byte piece: two pieces per part
    constexpr int PLANE = 64 * 64;              // bytes per plane in a stage buffer
    constexpr int SIDE = S * PLANE;
    constexpr int UB = S * 64 * 4;              // 16-byte units per side and stage
    constexpr int PER_T = 2 * UB / 256;
    static_assert(PER_T * 256 == 2 * UB && UB % kWave == 0 && PER_T == 2 * S,
                  "stage must split evenly, waves within a side, two units per digit round");
    static_assert(NBUF >= 2 && NBUF <= 4 && NBUF * 2 * SIDE <= kLdsBytes, "two to four stage buffers in LDS");

    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    const int wj = (wave / 2) * 32, wi = (wave % 2) * 32;
    // Fragment offsets within a plane: lane l reads row r = l % 32 of its
    // wave's 32 rows, 16-byte half h = l / 32 of the real (part 0) or
    // imaginary (part 1) 32 bytes.
    int fo_b[2], fo_a[2];
#pragma unroll
    for (int part = 0; part < 2; ++part) {
        const int h = lane / 32;
        const int rb = wj + lane % 32, ra = wi + lane % 32;
        fo_b[part] = rb * 64 + 16 * ((2 * part + h) ^ ((rb >> 2) & 3));
        fo_a[part] = ra * 64 + 16 * ((2 * part + h) ^ ((ra >> 2) & 3));
    }

    // Stage loads: unit u of a stage (u < UB: G' side) goes to byte 16 u of the
    // stage buffer; each wave instruction fills 1 KB, 16 bytes per lane.
    // issue_stage(st, q0, q1) issues this thread's units q0..q1-1 of the
    // PER_T = 2 S; the MFMA loop spreads them over its S digit rounds (a
    // burst would stall the wave on the memory pipeline with the MFMA unit
    // idle).
    const int npieces = kpad / KP;
    const unsigned char *gb = reinterpret_cast<const unsigned char *>(bd);
    const unsigned char *ga = reinterpret_cast<const unsigned char *>(ad);
    // Unit q of thread x: plane p = q % S of side q / S, row x / 4, slot x % 4
    // (the same row and slot for every q), so the source address is a per-
    // side lane offset plus uniform plane and stage strides.
    const int urow = threadIdx.x / 4, uunit = (threadIdx.x % 4) ^ ((urow >> 2) & 3);
    const size_t off_b = (((size_t)(uunit / 2) * npieces + uunit % 2) * ngp_pad + j0 + urow) * 16;
    const size_t off_a = (((size_t)(uunit / 2) * npieces + uunit % 2) * nc_pad + i0 + urow) * 16;
    const size_t plane_b = (size_t)2 * npieces * ngp_pad * 16, plane_a = (size_t)2 * npieces * nc_pad * 16;
    auto issue_stage = [&](int st, int q0, int q1) {
        unsigned char *buf = lds[st % NBUF] + 16 * (threadIdx.x & ~(kWave - 1));  // the wave's 1 KB
        const size_t kb = (size_t)st * (KS / KP) * 16;  // stage offset per column
#pragma unroll
        for (int q = q0; q < q1; ++q) {
            const unsigned char *src = q < S ? gb + (kb * ngp_pad + q * plane_b + off_b)
                                             : ga + (kb * nc_pad + (q - S) * plane_a + off_a);
            load_lds16(src, buf + 4096 * q);
        }
    };

    typedef OzAcc<D, S> A;
    acc_t acc_re[A::N], acc_im[A::N];
#pragma unroll
    for (int t = 0; t < A::N; ++t) {
        acc_re[t] = acc_t{};
        acc_im[t] = acc_t{};
    }
    // Row of result r of this lane (within the wave's 32 x 32) and its column.
    auto res_row = [&](int r) { return 8 * (r / 4) + 4 * (lane / 32) + r % 4; };
    const int ig = i0 + wi + lane % 32;
    auto fold = [&](int r, double &sr, double &sm) {
        sr = 0.0;
        sm = 0.0;
#pragma unroll
        for (int t = S - 1; t >= 0; --t) {
            double vr = (double)acc_re[A::first(t)][r], vm = (double)acc_im[A::first(t)][r];
            if (A::SPLIT && t == 3) {  // the level's second half (exact sum: both below 2^24)
                vr += (double)acc_re[A::SPLIT ? 4 : 0][r];
                vm += (double)acc_im[A::SPLIT ? 4 : 0][r];
            }
            sr = sr * A::INV_BASE + vr;
            sm = sm * A::INV_BASE + vm;
        }
    };
    // Column exponents of this lane's 16 results (eb, ea hold the padded
    // columns, so the clamped loads are all in bounds and issue together).
    auto load_exps = [&](int (&sc)[16]) {
        const int ea_i = ea[min(ig, nc_pad - 1)];
#pragma unroll
        for (int r = 0; r < 16; ++r) sc[r] = eb[min(j0 + wj + res_row(r), ngp_pad - 1)] + ea_i - 12;
    };
    // Chunked exact sums (FP16): running FP64 sums of the chunks kept,
    // unscaled, in the thread's own output elements of mout.
    auto flush = [&](bool first, bool last) {
        int sc[16];
        if (last) load_exps(sc);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int j = j0 + wj + res_row(r);
            if (j < ngpown && ig < ncouls) {
                double sr, sm;
                fold(r, sr, sm);
                cplx *o = &mout[(size_t)j * ncouls + ig];
                if (!first) {
                    const cplx prev = *o;
                    sr = prev.re + sr;
                    sm = prev.im + sm;
                }
                if (last) {
                    sr = ldexp(sr, sc[r]);
                    sm = ldexp(sm, sc[r]);
                }
                *o = cplx{sr, sm};
            }
        }
#pragma unroll
        for (int t = 0; t < A::N; ++t) {
            acc_re[t] = acc_t{};
            acc_im[t] = acc_t{};
        }
    };

    const int nst = kpad / KS;
    for (int st = 0; st < NBUF - 1 && st < nst; ++st) issue_stage(st, 0, PER_T);
    // Chunks of stages (CHUNK bands) with the flush between
    // them, outside the stage loop, so that the accumulators stay in place.
    constexpr bool CHUNKED = !T::ONE_CHUNK || LONG;
    constexpr int CST = CHUNKED ? T::CHUNK / KS : 1 << 30;
    bool drained = false;  // a flush waited for every outstanding load
    for (int c0 = 0; c0 < nst; c0 += CST) {
    const int c1 = min(nst, c0 + CST);
    for (int st = c0; st < c1; ++st) {
        // Wait for this wave's loads of stage st: the loads of the stages
        // issued after it (ahead of them, PER_T each) may stay in flight.
        const int ahead = drained ? 0 : min(NBUF - 2, nst - 1 - st);
        if (ahead <= 0)
            wait_vmcnt<0>();
        else if (ahead == 1)
            wait_vmcnt<PER_T>();
        else
            wait_vmcnt<(NBUF > 3 ? 2 * PER_T : PER_T)>();
        drained = false;
        // Stage st is in LDS for every wave, and every wave is done with stage
        // st - 1, whose buffer the loads of stage st + NBUF - 1 reuse.
        barrier_no_fence();
        const bool prefetch = st + NBUF - 1 < nst;
        const unsigned char *sB = lds[st % NBUF];
        const unsigned char *sA = sB + SIDE;
        frag fa[S][2];
#pragma unroll
        for (int q = 0; q < S; ++q)
#pragma unroll
            for (int part = 0; part < 2; ++part)
                fa[q][part] = *reinterpret_cast<const frag *>(sA + q * PLANE + fo_a[part]);
#pragma unroll
        for (int p = 0; p < S; ++p) {
            const frag br = *reinterpret_cast<const frag *>(sB + p * PLANE + fo_b[0]);
            const frag bi = *reinterpret_cast<const frag *>(sB + p * PLANE + fo_b[1]);
            const frag bn = T::neg(bi);
            // conj(b) a: re = br ar + bi ai, im = br ai - bi ar; the two
            // products into one accumulator are issued S - p MFMAs apart.
#pragma unroll
            for (int q = 0; q + p < S; ++q) {
                acc_re[A::idx(p, q)] = T::mfma32(br, fa[q][0], acc_re[A::idx(p, q)]);
                acc_im[A::idx(p, q)] = T::mfma32(br, fa[q][1], acc_im[A::idx(p, q)]);
            }
            if (prefetch) issue_stage(st + NBUF - 1, 2 * p, 2 * p + 2);
#pragma unroll
            for (int q = 0; q + p < S; ++q) {
                acc_re[A::idx(p, q)] = T::mfma32(bi, fa[q][1], acc_re[A::idx(p, q)]);
                acc_im[A::idx(p, q)] = T::mfma32(bn, fa[q][0], acc_im[A::idx(p, q)]);
            }
        }
    }
    if constexpr (CHUNKED) {
        flush(c0 == 0, c1 == nst);
        drained = true;
    }
    }

    if constexpr (!CHUNKED) {
        int sc[16];
        load_exps(sc);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int j = j0 + wj + res_row(r);
            if (j < ngpown && ig < ncouls) {
                double sr, sm;
                fold(r, sr, sm);
                // non-temporal: written once, read back once by the pole reduce
                __builtin_nontemporal_store(ldexp(sr, sc[r]), &mout[(size_t)j * ncouls + ig].re);
                __builtin_nontemporal_store(ldexp(sm, sc[r]), &mout[(size_t)j * ncouls + ig].im);
            }
        }
    }
}

// One tile of the layout-3 product as handed to gram_tiles32.
struct OzTile {
    const void *bd, *ad;
    const int *eb, *ea;
    int ngpown, ncouls, ngp_pad, nc_pad, kpad, j0, i0;
    cplx *mout;
};

// The tiles src.get(0), src.get(1), ... of a persistent block (get returns
// false past the last) as ONE stream of stages: stage g of the stream uses
// buffer g % NBUF and its loads are issued NBUF - 1 stages ahead, across tile
// boundaries, so a tile's first stages load while the previous tile
// finishes; otherwise as gram_tile32.
template <int S, typename D, int NBUF, bool LONG, class Src>
__device__ __forceinline__ void gram_tiles32(const Src &src, unsigned char (*lds)[Oz3<S>::STAGE])
{
    typedef OzDigit<D> T;
    typedef typename T::frag frag;
    typedef typename T::acc32 acc_t;
    constexpr int KS = T::KS32;
    constexpr int KP = 16 / (int)sizeof(D);
    constexpr int PLANE = 64 * 64;
    constexpr int SIDE = S * PLANE;
    constexpr int UB = S * 64 * 4;
    constexpr int PER_T = 2 * UB / 256;
    static_assert(PER_T * 256 == 2 * UB && UB % kWave == 0 && PER_T == 2 * S,
                  "stage must split evenly, waves within a side, two units per digit round");
    static_assert(NBUF >= 2 && NBUF <= 4 && NBUF * 2 * SIDE <= kLdsBytes, "two to four stage buffers in LDS");
    constexpr bool CHUNKED = !T::ONE_CHUNK || LONG;
    constexpr int CST = CHUNKED ? T::CHUNK / KS : 1 << 30;

    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    const int wj = (wave / 2) * 32, wi = (wave % 2) * 32;
    int fo_b[2], fo_a[2];
#pragma unroll
    for (int part = 0; part < 2; ++part) {
        const int h = lane / 32;
        const int rb = wj + lane % 32, ra = wi + lane % 32;
        fo_b[part] = rb * 64 + 16 * ((2 * part + h) ^ ((rb >> 2) & 3));
        fo_a[part] = ra * 64 + 16 * ((2 * part + h) ^ ((ra >> 2) & 3));
    }
    const int urow = threadIdx.x / 4, uunit = (threadIdx.x % 4) ^ ((urow >> 2) & 3);

    // Issue side: stage pst of tile pk (descriptor pt) is global stage gi.
    OzTile pt;
    int pk = 0, pst = 0, gi = 0, pnst = 0;
    bool pvalid = src.get(0, pt);
    if (!pvalid) return;
    const unsigned char *pgb = nullptr, *pga = nullptr;
    size_t poff_b = 0, poff_a = 0, pplane_b = 0, pplane_a = 0;
    auto set_issue_tile = [&]() {
        const int npieces = pt.kpad / KP;
        pnst = pt.kpad / KS;
        pgb = static_cast<const unsigned char *>(pt.bd);
        pga = static_cast<const unsigned char *>(pt.ad);
        poff_b = (((size_t)(uunit / 2) * npieces + uunit % 2) * pt.ngp_pad + pt.j0 + urow) * 16;
        poff_a = (((size_t)(uunit / 2) * npieces + uunit % 2) * pt.nc_pad + pt.i0 + urow) * 16;
        pplane_b = (size_t)2 * npieces * pt.ngp_pad * 16;
        pplane_a = (size_t)2 * npieces * pt.nc_pad * 16;
    };
    set_issue_tile();
    auto issue_units = [&](int q0, int q1) {
        unsigned char *buf = lds[gi % NBUF] + 16 * (threadIdx.x & ~(kWave - 1));
        const size_t kb = (size_t)pst * (KS / KP) * 16;
#pragma unroll
        for (int q = q0; q < q1; ++q) {
            const unsigned char *s = q < S ? pgb + (kb * pt.ngp_pad + q * pplane_b + poff_b)
                                           : pga + (kb * pt.nc_pad + (q - S) * pplane_a + poff_a);
            load_lds16(s, buf + 4096 * q);
        }
    };
    auto advance_issue = [&]() {
        ++gi;
        if (++pst == pnst) {
            pst = 0;
            ++pk;
            pvalid = src.get(pk, pt);
            if (pvalid) set_issue_tile();
        }
    };
    for (int i = 0; i < NBUF - 1 && pvalid; ++i) {
        issue_units(0, PER_T);
        advance_issue();
    }

    // Compute side: tile ck (descriptor ct), global stage g.
    OzTile ct;
    (void)src.get(0, ct);
    int ck = 0, g = 0;
    typedef OzAcc<D, S> A;
    acc_t acc_re[A::N], acc_im[A::N];
#pragma unroll
    for (int t = 0; t < A::N; ++t) {
        acc_re[t] = acc_t{};
        acc_im[t] = acc_t{};
    }
    auto res_row = [&](int r) { return 8 * (r / 4) + 4 * (lane / 32) + r % 4; };
    auto fold = [&](int r, double &sr, double &sm) {
        sr = 0.0;
        sm = 0.0;
#pragma unroll
        for (int t = S - 1; t >= 0; --t) {
            double vr = (double)acc_re[A::first(t)][r], vm = (double)acc_im[A::first(t)][r];
            if (A::SPLIT && t == 3) {  // the level's second half (exact sum: both below 2^24)
                vr += (double)acc_re[A::SPLIT ? 4 : 0][r];
                vm += (double)acc_im[A::SPLIT ? 4 : 0][r];
            }
            sr = sr * A::INV_BASE + vr;
            sm = sm * A::INV_BASE + vm;
        }
    };
    // first: the tile's first chunk (store); last: its last (scale).  Clears
    // the accumulators.
    // Column exponents of this lane's 16 results (clamped into the padded
    // exponent arrays, so the loads issue together).
    auto load_exps = [&](int (&sc)[16]) {
        const int ea_i = ct.ea[min(ct.i0 + wi + lane % 32, ct.nc_pad - 1)];
#pragma unroll
        for (int r = 0; r < 16; ++r) sc[r] = ct.eb[min(ct.j0 + wj + res_row(r), ct.ngp_pad - 1)] + ea_i - 12;
    };
    auto flush = [&](bool first, bool last) {
        const int ig = ct.i0 + wi + lane % 32;
        int sc[16];
        if (last) load_exps(sc);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int j = ct.j0 + wj + res_row(r);
            if (j < ct.ngpown && ig < ct.ncouls) {
                double sr, sm;
                fold(r, sr, sm);
                cplx *o = &ct.mout[(size_t)j * ct.ncouls + ig];
                if (!first) {
                    const cplx prev = *o;
                    sr = prev.re + sr;
                    sm = prev.im + sm;
                }
                if (last) {
                    sr = ldexp(sr, sc[r]);
                    sm = ldexp(sm, sc[r]);
                }
                *o = cplx{sr, sm};
            }
        }
#pragma unroll
        for (int t = 0; t < A::N; ++t) {
            acc_re[t] = acc_t{};
            acc_im[t] = acc_t{};
        }
    };
    bool drained = false;  // memory traffic since the last wait: wait for everything
    for (;;) {
        const int nst = ct.kpad / KS;
        for (int c0 = 0; c0 < nst; c0 += CST) {
            const int c1 = min(nst, c0 + CST);
            for (int st = c0; st < c1; ++st, ++g) {
                // Stages issued after g (PER_T loads each) may stay in flight.
                const int ahead = drained ? 0 : gi - 1 - g;
                if (ahead <= 0)
                    wait_vmcnt<0>();
                else if (ahead == 1)
                    wait_vmcnt<PER_T>();
                else
                    wait_vmcnt<(NBUF > 3 ? 2 * PER_T : PER_T)>();
                drained = false;
                // Stage g is in LDS for every wave, and every wave is done
                // with stage g - 1, whose buffer stage g + NBUF - 1 reuses.
                barrier_no_fence();
                const bool prefetch = pvalid && gi < g + NBUF;
                const unsigned char *sB = lds[g % NBUF];
                const unsigned char *sA = sB + SIDE;
                frag fa[S][2];
#pragma unroll
                for (int q = 0; q < S; ++q)
#pragma unroll
                    for (int part = 0; part < 2; ++part)
                        fa[q][part] = *reinterpret_cast<const frag *>(sA + q * PLANE + fo_a[part]);
#pragma unroll
                for (int p = 0; p < S; ++p) {
                    const frag br = *reinterpret_cast<const frag *>(sB + p * PLANE + fo_b[0]);
                    const frag bi = *reinterpret_cast<const frag *>(sB + p * PLANE + fo_b[1]);
                    const frag bn = T::neg(bi);
#pragma unroll
                    for (int q = 0; q + p < S; ++q) {
                        acc_re[A::idx(p, q)] = T::mfma32(br, fa[q][0], acc_re[A::idx(p, q)]);
                        acc_im[A::idx(p, q)] = T::mfma32(br, fa[q][1], acc_im[A::idx(p, q)]);
                    }
                    if (prefetch) issue_units(2 * p, 2 * p + 2);
#pragma unroll
                    for (int q = 0; q + p < S; ++q) {
                        acc_re[A::idx(p, q)] = T::mfma32(bi, fa[q][1], acc_re[A::idx(p, q)]);
                        acc_im[A::idx(p, q)] = T::mfma32(bn, fa[q][0], acc_im[A::idx(p, q)]);
                    }
                }
                if (prefetch) advance_issue();
            }
            if constexpr (CHUNKED) {
                flush(c0 == 0, c1 == nst);
                drained = true;
            }
        }
        if constexpr (!CHUNKED) {
            const int ig = ct.i0 + wi + lane % 32;
            int sc[16];
            load_exps(sc);
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int j = ct.j0 + wj + res_row(r);
                if (j < ct.ngpown && ig < ct.ncouls) {
                    double sr, sm;
                    fold(r, sr, sm);
                    // non-temporal: written once, read back once by the pole
                    // reduce (C1 INT8 step 2.64 -> 2.60 ms, 5 slices 2.20 -> 2.16)
                    __builtin_nontemporal_store(ldexp(sr, sc[r]), &ct.mout[(size_t)j * ct.ncouls + ig].re);
                    __builtin_nontemporal_store(ldexp(sm, sc[r]), &ct.mout[(size_t)j * ct.ncouls + ig].im);
                }
            }
#pragma unroll
            for (int t = 0; t < A::N; ++t) {
                acc_re[t] = acc_t{};
                acc_im[t] = acc_t{};
            }
            drained = true;
        }
        if (!src.get(++ck, ct)) break;
    }
}

struct OzLayout {
    size_t eb, ea, bd, ad, total;
    int ngp_pad, nc_pad, kpad, K;
};

inline OzLayout oz_layout(const DevProblem &p, int slices, bool int8)
{
    const int ks = int8 ? OzDigit<int8_t>::KS : OzDigit<_Float16>::KS;
    const size_t dsz = int8 ? 1 : 2;
    OzLayout L;
    L.K = p.nb - p.nv;
    L.kpad = round_up(L.K > 0 ? L.K : 1, ks);
    L.ngp_pad = round_up(p.ngpown, kOzTile);
    L.nc_pad = round_up(p.ncouls, kOzTile);
    size_t off = 0;
    L.eb = off; off += align256(sizeof(int) * L.ngp_pad);
    L.ea = off; off += align256(sizeof(int) * L.nc_pad);
    L.bd = off; off += align256(dsz * (size_t)slices * 2 * L.ngp_pad * L.kpad);
    L.ad = off; off += align256(dsz * (size_t)slices * 2 * L.nc_pad * L.kpad);
    L.total = off;
    return L;
}

// G' tiles per tile group: the given count, or (0) as many tiles as fill about
// 2.5 MB of an XCD's 4 MB L2 with their digit planes (tile_bytes each).
inline int tile_group(int group, int ntj, size_t tile_bytes)
{
    if (group <= 0) group = (int)((size_t)(5 << 19) / tile_bytes);
    return group < 1 ? 1 : group > ntj ? ntj : group;
}

}  // namespace oz
}  // namespace gpp
