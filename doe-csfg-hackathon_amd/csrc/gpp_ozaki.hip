// gpp_ozaki.hip -- the complex Gram product of the factored form as an
// Ozaki-scheme split into FP16 slices, multiplied on FP16 MFMA with exact FP32
// accumulation, recombined in FP64.
//
//   Mu[j][ig] = sum_{k in [nv, nb)} conj(Bg[k][j]) * aqsn[k][ig]
//
// Splitting.  Every column c of an operand (a G' column j of Bg, a G-vector
// ig of aqsn) gets a power-of-two scale 2^e_c > max_k |re|, |im|.  Each real
// part x is written as
//   x = 2^(e_c - 6) * sum_{p<S} d_p B^-p + r,   d_0 in [-64, 64],
//   d_p in [-B/2, B/2] (p >= 1),   |r| <= 2^(e_c - 7) B^(1-S)
// (d_0 = rint(x 2^(6-e_c)), each later digit rounds B times the previous
// remainder).  B = 254: digits of about 8 bits after the first, exact in FP16
// and int8, 39 bits at S = 5 and 47 at S = 6; FP16 at S = 6 uses B = 128 (42
// bits; OzAcc in gpp_ozaki_dev.hpp).  The product of digits p and q has weight
// 2^(e_j + e_ig - 12) B^-(p+q); the pairs of one level t = p+q share an
// accumulator, and pairs with t >= S are dropped (their weight is below the
// split's own truncation).  Levels fold to FP64 by Horner in 1/B.
//
// Accumulation.  A level-t accumulator receives (t+1) pairs x 2 real products
// (conj(b)a: br ar + bi ai, br ai - bi ar) per band.  FP16 digits accumulate
// in FP32 over chunks of 256 bands, exactly (integers below 2^24, which FP32
// holds whatever order the MFMA adds in) for every level with B = 128 and for
// levels 0-3 with B = 254 (level 3 in two accumulators); levels 4 and up of
// the latter round, within 2^-40 of the chunk's largest possible sum.  After
// each chunk the levels are converted to FP64 and accumulated.
//
// The same kernel runs with INT8 digits (v_mfma_i32_16x16x64_i8): INT32 sums
// are exact over chunks of 8192 bands (at most 161544 per band and level at
// S = 6), so up to that band count the levels are folded to FP64 once, after
// the whole band range, and longer ranges are flushed per chunk; MI355X runs
// INT8 MFMA at twice the FP16 rate and the staged digits are half the bytes.
//
// MFMA lane layouts (measured on gfx950), 16 bytes of operand per lane:
//   f32_16x16x32_f16: A[l%16][8(l/16)+e], B[8(l/16)+e][l%16]
//   i32_16x16x64_i8 : A[l%16][16(l/16)+e], B[16(l/16)+e][l%16]
//   both            : result r of lane l = C[4(l/16)+r][l%16].
//   f32_32x32x16_f16: A[l%32][8(l/32)+e], B[8(l/32)+e][l%32]
//   i32_32x32x32_i8 : A[l%32][16(l/32)+e], B[16(l/32)+e][l%32]
//   both            : result r of lane l = C[8(r/4)+4(l/32)+r%4][l%32].
#include "gpp_ozaki_dev.hpp"

namespace gpp {
namespace {

using namespace oz;

// LDS: per digit plane p and tile row (a column j or ig), one 128-byte row
// holding the stage's real-part digits (16-byte units 0-3) and imaginary-part
// digits (units 4-7); logical unit u of row r sits in slot u ^ (r & 7), which
// makes both the contiguous stores and the 16-row MFMA fragment reads free of
// LDS bank conflicts on gfx950 (a padded 80-byte pitch is not).
constexpr int kOzRowBytes = 128;

// Block layouts: NWJ x NWI waves, each owning SJ x SI sub-tiles of 16 x 16,
// so a block tile is TJ = 16 NWJ SJ G' columns by TI = 16 NWI SI G-vectors.
//   layout 0: 64 x 64, 2 x 2 waves of 32 x 32 (one wave per SIMD, 512
//             registers each), stages staged through registers
//   layout 1: 64 x 64, 2 x 4 waves of 32 x 16 (two waves per SIMD), staged
//             through registers
//   layout 2: 64 x 32, 2 x 2 waves of 32 x 16, stages loaded straight from
//             memory into LDS (global_load_lds_dwordx4, no staging registers);
//             two blocks share a CU, so one block's loads and barriers overlap
//             the other block's MFMAs.
template <int L> struct OzWaves;
template <> struct OzWaves<0> {
    static constexpr int NWJ = 2, NWI = 2, SJ = 2, SI = 2, BLOCKS_PER_CU = 1;
    static constexpr bool DMA = false;
};
template <> struct OzWaves<1> {
    static constexpr int NWJ = 2, NWI = 4, SJ = 2, SI = 1, BLOCKS_PER_CU = 1;
    static constexpr bool DMA = false;
};
template <> struct OzWaves<2> {
    static constexpr int NWJ = 2, NWI = 2, SJ = 2, SI = 1, BLOCKS_PER_CU = 2;
    static constexpr bool DMA = true;
};

// Column exponents: grid ncols_pad / 64.
__global__ __launch_bounds__(256) void k_oz_colexp(const cplx *__restrict__ src, int ld, int k_begin, int k_end,
                                                   int ncols, int ncols_pad, int *__restrict__ exps)
{
    __shared__ double part[4][64];
    colexp_block(src, ld, k_begin, k_end, ncols, ncols_pad, exps, blockIdx.x, part);
}

// Digit planes: grid (ncols_pad / 256, kpad / KP).
template <int S, typename D>
__global__ __launch_bounds__(256) void k_oz_split(const cplx *__restrict__ src, int ld, int k_begin, int K,
                                                  int ncols, int ncols_pad, int kpad,
                                                  const int *__restrict__ exps, D *__restrict__ out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < ncols_pad) split_piece<S, D>(src, ld, k_begin, K, ncols, ncols_pad, kpad, exps, out, c, blockIdx.y);
}

// Mu tiles from the digit planes.  1-D grid of (ngpown_pad/TJ) x (ncouls_pad/TI)
// tiles, 64*NWJ*NWI threads.  Tile order: groups of gj G' tiles; within a
// group, G-vector tiles in turn with the group's G' tiles fastest; the order
// is split into one contiguous range per XCD.  The blocks of an XCD then
// share the digit planes of a few G' tiles (kept in its L2) and stream the
// G-vector tiles once per group.
template <int S, typename D, int LAYOUT>
__global__ __launch_bounds__(64 * OzWaves<LAYOUT>::NWJ * OzWaves<LAYOUT>::NWI, OzWaves<LAYOUT>::BLOCKS_PER_CU)
void k_gram_ozaki(const D *__restrict__ bd, const D *__restrict__ ad, const int *__restrict__ eb,
                  const int *__restrict__ ea, int ngpown, int ncouls, int ngp_pad, int nc_pad, int kpad, int gj,
                  cplx *__restrict__ mout)
{
    typedef OzDigit<D> T;
    typedef OzWaves<LAYOUT> W;
    constexpr int SJ = W::SJ, SI = W::SI, NTH = 64 * W::NWJ * W::NWI;
    constexpr int TJ = 16 * W::NWJ * SJ, TI = 16 * W::NWI * SI;
    typedef typename T::frag frag;
    typedef typename T::acc acc_t;
    constexpr int KS = T::KS;
    __shared__ __attribute__((aligned(16))) unsigned char sB[S][TJ][kOzRowBytes];
    __shared__ __attribute__((aligned(16))) unsigned char sA[S][TI][kOzRowBytes];
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    const int wj = (wave / W::NWI) * (16 * SJ), wi = (wave % W::NWI) * (16 * SI);
    int j0, i0;
    {
        const int ntj = ngp_pad / TJ, nti = nc_pad / TI;
        const int t = xcd_tile(blockIdx.x, ntj * nti);
        const int g = t / (gj * nti), rem = t - g * gj * nti;
        const int gsz = min(gj, ntj - g * gj);
        j0 = (g * gj + rem % gsz) * TJ;
        i0 = (rem / gsz) * TI;
    }
    // Fragment slots: every row this lane reads is = lane (mod 8).
    const int fo_re = 16 * ((lane / 16) ^ (lane & 7));
    const int fo_im = 16 * ((4 + lane / 16) ^ (lane & 7));

    // Level accumulators.  With INT8 the levels are exact over the whole band
    // range and are folded to FP64 once, in the epilogue; with chunked exact
    // sums (FP16) the FP64 running sums of the chunks are kept, unscaled, in
    // the thread's own output elements of mout (registers are the limit here)
    // and scaled after the last chunk.
    typedef OzAcc<D, S> A;
    acc_t acc_re[A::N][SJ][SI], acc_im[A::N][SJ][SI];
#pragma unroll
    for (int sj = 0; sj < SJ; ++sj)
#pragma unroll
        for (int si = 0; si < SI; ++si)
#pragma unroll
            for (int t = 0; t < A::N; ++t) acc_re[t][sj][si] = acc_im[t][sj][si] = acc_t{0, 0, 0, 0};
    // Horner over the levels (smallest first) of result r of sub-tile (sj, si).
    auto fold = [&](int sj, int si, int r, double &sr, double &sm) {
        sr = 0.0;
        sm = 0.0;
#pragma unroll
        for (int t = S - 1; t >= 0; --t) {
            double vr = (double)acc_re[A::first(t)][sj][si][r], vm = (double)acc_im[A::first(t)][sj][si][r];
            if (A::SPLIT && t == 3) {  // the level's second half (exact sum: both below 2^24)
                vr += (double)acc_re[A::SPLIT ? 4 : 0][sj][si][r];
                vm += (double)acc_im[A::SPLIT ? 4 : 0][sj][si][r];
            }
            sr = sr * A::INV_BASE + vr;
            sm = sm * A::INV_BASE + vm;
        }
    };

    // Fold the level accumulators into the FP64 sums (chunked form only):
    // first chunk stores, later chunks add, the last one also scales.
    auto flush = [&](bool first, bool last) {
#pragma unroll
        for (int sj = 0; sj < SJ; ++sj)
#pragma unroll
            for (int si = 0; si < SI; ++si)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int j = j0 + wj + 16 * sj + 4 * (lane / 16) + r;
                    const int ig = i0 + wi + 16 * si + lane % 16;
                    if (j < ngpown && ig < ncouls) {
                        double sr, sm;
                        fold(sj, si, r, sr, sm);
                        cplx *o = &mout[(size_t)j * ncouls + ig];
                        if (!first) {
                            const cplx prev = *o;
                            sr = prev.re + sr;
                            sm = prev.im + sm;
                        }
                        if (last) {
                            const int sc = eb[j] + ea[ig] - 12;
                            sr = ldexp(sr, sc);
                            sm = ldexp(sm, sc);
                        }
                        *o = cplx{sr, sm};
                    }
                }
#pragma unroll
        for (int sj = 0; sj < SJ; ++sj)
#pragma unroll
            for (int si = 0; si < SI; ++si)
#pragma unroll
                for (int t = 0; t < A::N; ++t) acc_re[t][sj][si] = acc_im[t][sj][si] = acc_t{0, 0, 0, 0};
    };

    // Digit planes in HBM: 16-byte pieces (KP bands of one column), pieces of
    // consecutive columns contiguous, planes of kpad/KP piece rows.
    constexpr int KP = 16 / (int)sizeof(D);
    const int npieces = kpad / KP;
    const unsigned char *gb = reinterpret_cast<const unsigned char *>(bd);
    const unsigned char *ga = reinterpret_cast<const unsigned char *>(ad);
    // 16-byte units per stage: S planes x TJ rows x 8 (G' side), then S planes
    // x TI rows x 8 (G-vector side); unit v of a side lands at byte 16 v of its
    // LDS array ([plane][row][slot]).
    constexpr int UB = S * TJ * 8, UA = S * TI * 8;
    constexpr int PER_T = (UB + UA) / NTH;
    static_assert(PER_T * NTH == UB + UA && UB % kWave == 0, "stage must split evenly, waves within a side");
    // Global address of unit u of the stage starting at piece kq0.
    auto unit_src = [&](int u, int kq0) -> const unsigned char * {
        const bool side = u >= UB;
        const int v = side ? u - UB : u;
        const int rows = side ? TI : TJ;
        const int slot = v % 8, row = (v / 8) % rows, p = v / (8 * rows);
        const int lu = slot ^ (row & 7);  // logical unit: part = lu / 4, 16-byte piece lu % 4
        const int npad = side ? nc_pad : ngp_pad;
        const int col = (side ? i0 : j0) + row;
        return (side ? ga : gb) + (((size_t)(p * 2 + lu / 4) * npieces + kq0 + lu % 4) * npad + col) * 16;
    };
    // Register staging (layouts 0, 1): the next stage's units are loaded into
    // registers while the current stage computes, then stored to LDS between
    // two barriers.
    v4i stage_reg[W::DMA ? 1 : PER_T];
    auto load_stage = [&](int k0) {
#pragma unroll
        for (int q = 0; q < PER_T; ++q)
            stage_reg[q] = *reinterpret_cast<const v4i *>(unit_src(threadIdx.x + q * NTH, k0 / KP));
    };
    auto store_stage = [&]() {
#pragma unroll
        for (int q = 0; q < PER_T; ++q) {
            const int u = threadIdx.x + q * NTH;
            unsigned char *dst = u >= UB ? &sA[0][0][0] + 16 * (u - UB) : &sB[0][0][0] + 16 * u;
            *reinterpret_cast<v4i *>(dst) = stage_reg[q];
        }
    };
    // Direct loads (layout 2): each wave instruction fills 1 KB of LDS, 16
    // bytes per lane in lane order, from the addresses its lanes give.
    auto dma_stage = [&](int k0) {
#pragma unroll
        for (int q = 0; q < PER_T; ++q) {
            const int u = threadIdx.x + q * NTH;
            const int ub = (threadIdx.x & ~(kWave - 1)) + q * NTH;  // the wave's first unit
            unsigned char *dst = ub >= UB ? &sA[0][0][0] + 16 * (ub - UB) : &sB[0][0][0] + 16 * ub;
            load_lds16(unit_src(u, k0 / KP), dst);
        }
    };
    if constexpr (W::DMA) {
        dma_stage(0);
        __builtin_amdgcn_s_waitcnt(0);
    } else {
        load_stage(0);
        store_stage();
    }
    __syncthreads();
    for (int k0 = 0; k0 < kpad; k0 += KS) {
        if constexpr (!W::DMA) {
            if (k0 + KS < kpad) load_stage(k0 + KS);
        }
        // G' side fragments of digit p are held across its pairs; the G-vector
        // side fragments of digit q are read from LDS for each pair (fewer
        // live registers, which leaves room for the next stage's loads).
#pragma unroll
        for (int p = 0; p < S; ++p) {
            frag br[SJ], bi[SJ], bn[SJ];
#pragma unroll
            for (int sj = 0; sj < SJ; ++sj) {
                br[sj] = *reinterpret_cast<const frag *>(&sB[p][wj + 16 * sj + lane % 16][fo_re]);
                bi[sj] = *reinterpret_cast<const frag *>(&sB[p][wj + 16 * sj + lane % 16][fo_im]);
                bn[sj] = T::neg(bi[sj]);
            }
#pragma unroll
            for (int q = 0; q + p < S; ++q) {
                const int t = A::idx(p, q);
                frag fa[2][SI];  // [real, imaginary part][sub-tile] of digit q
#pragma unroll
                for (int part = 0; part < 2; ++part)
#pragma unroll
                    for (int si = 0; si < SI; ++si)
                        fa[part][si] = *reinterpret_cast<const frag *>(
                            &sA[q][wi + 16 * si + lane % 16][part ? fo_im : fo_re]);
                // conj(b) * a: re = br ar + bi ai, im = br ai - bi ar.  Each group of
                // four MFMAs writes four different accumulators.
#pragma unroll
                for (int sj = 0; sj < SJ; ++sj)
#pragma unroll
                    for (int si = 0; si < SI; ++si) acc_re[t][sj][si] = T::mfma(br[sj], fa[0][si], acc_re[t][sj][si]);
#pragma unroll
                for (int sj = 0; sj < SJ; ++sj)
#pragma unroll
                    for (int si = 0; si < SI; ++si) acc_im[t][sj][si] = T::mfma(br[sj], fa[1][si], acc_im[t][sj][si]);
#pragma unroll
                for (int sj = 0; sj < SJ; ++sj)
#pragma unroll
                    for (int si = 0; si < SI; ++si) acc_re[t][sj][si] = T::mfma(bi[sj], fa[1][si], acc_re[t][sj][si]);
#pragma unroll
                for (int sj = 0; sj < SJ; ++sj)
#pragma unroll
                    for (int si = 0; si < SI; ++si) acc_im[t][sj][si] = T::mfma(bn[sj], fa[0][si], acc_im[t][sj][si]);
            }
        }
        if constexpr (W::DMA) {
            // All waves are done with the stage before it is overwritten; the
            // flush's memory traffic overlaps the loads.
            if (k0 + KS < kpad) {
                __syncthreads();
                dma_stage(k0 + KS);
            }
            if constexpr (!T::ONE_CHUNK) {
                if ((k0 + KS) % T::CHUNK == 0 || k0 + KS >= kpad) flush(k0 + KS <= T::CHUNK, k0 + KS >= kpad);
            }
            if (k0 + KS < kpad) {
                __builtin_amdgcn_s_waitcnt(0);
                __syncthreads();
            }
        } else {
            __syncthreads();
            if (k0 + KS < kpad) store_stage();  // the staged registers die before the flush
            if constexpr (!T::ONE_CHUNK) {
                if ((k0 + KS) % T::CHUNK == 0 || k0 + KS >= kpad) flush(k0 + KS <= T::CHUNK, k0 + KS >= kpad);
            }
            if (k0 + KS < kpad) __syncthreads();
        }
    }

    if constexpr (T::ONE_CHUNK) {
#pragma unroll
        for (int sj = 0; sj < SJ; ++sj)
#pragma unroll
            for (int si = 0; si < SI; ++si)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int j = j0 + wj + 16 * sj + 4 * (lane / 16) + r;
                    const int ig = i0 + wi + 16 * si + lane % 16;
                    if (j < ngpown && ig < ncouls) {
                        const int sc = eb[j] + ea[ig] - 12;
                        double sr, sm;
                        fold(sj, si, r, sr, sm);
                        mout[(size_t)j * ncouls + ig] = cplx{ldexp(sr, sc), ldexp(sm, sc)};
                    }
                }
    }
}

template <int S, typename D, int NBUF, bool LONG>
__global__ __launch_bounds__(256, 1) void k_gram_ozaki32(const D *__restrict__ bd, const D *__restrict__ ad,
                                                         const int *__restrict__ eb, const int *__restrict__ ea,
                                                         int ngpown, int ncouls, int ngp_pad, int nc_pad, int kpad,
                                                         int gj, cplx *__restrict__ mout)
{
    __shared__ __attribute__((aligned(16))) unsigned char lds[NBUF][Oz3<S>::STAGE];
    const int ntj = ngp_pad / kOz3Tile, nti = nc_pad / kOz3Tile;
    const int t = xcd_tile(blockIdx.x, ntj * nti);
    const int g = t / (gj * nti), rem = t - g * gj * nti;
    const int gsz = min(gj, ntj - g * gj);
    gram_tile32<S, D, NBUF, LONG>(bd, ad, eb, ea, ngpown, ncouls, ngp_pad, nc_pad, kpad, (g * gj + rem % gsz) * kOz3Tile,
                            (rem / gsz) * kOz3Tile, mout, lds);
}

// Layout 4: the layout-3 tiles in persistent blocks (one per CU), the block's
// tiles taken in the same grouped, XCD-ordered sequence with a stride of the
// grid, as one stream of stages (the next tile's first stages load while the
// current one finishes).
template <typename D> struct OzTileSeq {
    const D *bd, *ad;
    const int *eb, *ea;
    int ngpown, ncouls, ngp_pad, nc_pad, kpad, gj;
    cplx *mout;
    __device__ bool get(int k, OzTile &t) const
    {
        const int ntj = ngp_pad / kOz3Tile, nti = nc_pad / kOz3Tile, nt = ntj * nti;
        const long long b = (long long)blockIdx.x + (long long)k * gridDim.x;
        if (b >= nt) return false;
        const int tt = xcd_tile((int)b, nt);
        const int g = tt / (gj * nti), rem = tt - g * gj * nti;
        const int gsz = min(gj, ntj - g * gj);
        t = OzTile{bd, ad, eb, ea, ngpown, ncouls, ngp_pad, nc_pad, kpad, (g * gj + rem % gsz) * kOz3Tile,
                   (rem / gsz) * kOz3Tile, mout};
        return true;
    }
};

template <int S, typename D, int NBUF, bool LONG>
__global__ __launch_bounds__(256, 1) void k_gram_ozaki32p(OzTileSeq<D> seq)
{
    __shared__ __attribute__((aligned(16))) unsigned char lds[NBUF][Oz3<S>::STAGE];
    gram_tiles32<S, D, NBUF, LONG>(seq, lds);
}

// Compute units of the current device (cached per device).
int device_cus()
{
    static int cache[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (!cache[dev]) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
        cache[dev] = n;
    }
    return cache[dev];
}

template <int S, typename D, int LAYOUT>
void launch_gram_tiles(const OzLayout &L, const DevProblem &p, const D *bd, const D *ad, const int *eb, const int *ea,
                       int group, cplx *mout, hipStream_t s)
{
    typedef OzWaves<LAYOUT> W;
    constexpr int TJ = 16 * W::NWJ * W::SJ, TI = 16 * W::NWI * W::SI;
    const int ntj = L.ngp_pad / TJ;
    const dim3 grid((unsigned)ntj * (unsigned)(L.nc_pad / TI));
    hipLaunchKernelGGL((k_gram_ozaki<S, D, LAYOUT>), grid, dim3(64 * W::NWJ * W::NWI), 0, s, bd, ad, eb, ea, p.ngpown,
                       p.ncouls, L.ngp_pad, L.nc_pad, L.kpad,
                       tile_group(group, ntj, (size_t)S * 2 * TJ * L.kpad * sizeof(D)), mout);
}

template <int S, typename D>
hipError_t launch_ozaki_s(const DevProblem &p, int layout, int group, const cplx *bg, void *work, cplx *mout,
                          hipStream_t s)
{
    const OzLayout L = oz_layout(p, S, sizeof(D) == 1);
    char *w = static_cast<char *>(work);
    int *eb = reinterpret_cast<int *>(w + L.eb);
    int *ea = reinterpret_cast<int *>(w + L.ea);
    D *bd = reinterpret_cast<D *>(w + L.bd);
    D *ad = reinterpret_cast<D *>(w + L.ad);
    constexpr int KP = 16 / (int)sizeof(D);
    hipLaunchKernelGGL(k_oz_colexp, dim3(L.ngp_pad / 64), dim3(256), 0, s, bg, p.ngpown, p.nv, p.nb, p.ngpown,
                       L.ngp_pad, eb);
    hipLaunchKernelGGL(k_oz_colexp, dim3(L.nc_pad / 64), dim3(256), 0, s, p.aqsn, p.ncouls, p.nv, p.nb, p.ncouls,
                       L.nc_pad, ea);
    hipLaunchKernelGGL((k_oz_split<S, D>), dim3((L.ngp_pad + 255) / 256, L.kpad / KP), dim3(256), 0, s, bg,
                       p.ngpown, p.nv, L.K, p.ngpown, L.ngp_pad, L.kpad, eb, bd);
    hipLaunchKernelGGL((k_oz_split<S, D>), dim3((L.nc_pad + 255) / 256, L.kpad / KP), dim3(256), 0, s, p.aqsn,
                       p.ncouls, p.nv, L.K, p.ncouls, L.nc_pad, L.kpad, ea, ad);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // INT8 band ranges beyond one exact chunk run the layout-3 tile that
    // flushes per chunk, whatever the layout option.
    const bool long_range = OzDigit<D>::ONE_CHUNK && L.kpad > OzDigit<D>::CHUNK;
    if (layout == 4) {  // persistent layout 3
        const int ntj = L.ngp_pad / kOz3Tile;
        const unsigned nt = (unsigned)ntj * (unsigned)(L.nc_pad / kOz3Tile);
        const unsigned cus = (unsigned)device_cus();
        const OzTileSeq<D> seq{bd, ad, eb, ea, p.ngpown, p.ncouls, L.ngp_pad, L.nc_pad, L.kpad,
                               tile_group(group, ntj, (size_t)S * 2 * kOz3Tile * L.kpad * sizeof(D)), mout};
        const dim3 grid(nt < cus ? nt : cus);
        if (long_range)
            hipLaunchKernelGGL((k_gram_ozaki32p<S, D, oz3_buffers(S), true>), grid, dim3(256), 0, s, seq);
        else
            hipLaunchKernelGGL((k_gram_ozaki32p<S, D, oz3_buffers(S), false>), grid, dim3(256), 0, s, seq);
        return hipGetLastError();
    }
    switch (long_range ? 3 : layout) {
    case 0: launch_gram_tiles<S, D, 0>(L, p, bd, ad, eb, ea, group, mout, s); break;
    case 1: launch_gram_tiles<S, D, 1>(L, p, bd, ad, eb, ea, group, mout, s); break;
    case 2: launch_gram_tiles<S, D, 2>(L, p, bd, ad, eb, ea, group, mout, s); break;
    default: {
        const int ntj = L.ngp_pad / kOz3Tile;
        const dim3 grid((unsigned)ntj * (unsigned)(L.nc_pad / kOz3Tile));
        const int gj = tile_group(group, ntj, (size_t)S * 2 * kOz3Tile * L.kpad * sizeof(D));
        if (long_range)
            hipLaunchKernelGGL((k_gram_ozaki32<S, D, oz3_buffers(S), true>), grid, dim3(256), 0, s, bd, ad, eb, ea,
                               p.ngpown, p.ncouls, L.ngp_pad, L.nc_pad, L.kpad, gj, mout);
        else
            hipLaunchKernelGGL((k_gram_ozaki32<S, D, oz3_buffers(S), false>), grid, dim3(256), 0, s, bd, ad, eb, ea,
                               p.ngpown, p.ncouls, L.ngp_pad, L.nc_pad, L.kpad, gj, mout);
        break;
    }
    }
    return hipGetLastError();
}

template <typename D>
hipError_t launch_ozaki_d(const DevProblem &p, int slices, int layout, int group, const cplx *bg, void *work,
                          cplx *mout, hipStream_t s)
{
    switch (slices) {
    case 3: return launch_ozaki_s<3, D>(p, layout, group, bg, work, mout, s);
    case 4: return launch_ozaki_s<4, D>(p, layout, group, bg, work, mout, s);
    case 5: return launch_ozaki_s<5, D>(p, layout, group, bg, work, mout, s);
    case 6: return launch_ozaki_s<6, D>(p, layout, group, bg, work, mout, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

size_t ozaki_workspace_bytes(const DevProblem &p, int slices, bool int8) { return oz_layout(p, slices, int8).total; }

hipError_t launch_gram_ozaki(const DevProblem &p, int slices, bool int8, int layout, int group, const cplx *bg,
                             void *work, cplx *mout, hipStream_t s)
{
    return int8 ? launch_ozaki_d<int8_t>(p, slices, layout, group, bg, work, mout, s)
                : launch_ozaki_d<_Float16>(p, slices, layout, group, bg, work, mout, s);
}

}  // namespace gpp
