// gpp_factored_dev.hpp -- device building blocks of the factored form, shared
// by the single-instance kernels (gpp_factored.hip, one grid per stage) and
// the batched kernels (gpp_batch.hip, one work list per stage).
//
// Factorisation (wx_array is band-independent, gppComplex.cpp:323-329):
//   Mu[j,ig] = sum_{n1 >= nv} conj(aqsm[n1,igp(j)]) * aqsn[n1,ig]
//   Mo[j,ig] = sum_{n1 <  nv} conj(aqsm[n1,igp(j)]) * aqsn[n1,ig]
//   achtemp[iw] = sum_j 0.5 vcoul[igp(j)] sum_ig (P*Mu + sch*Mo)
//   asxtemp[iw] = sum_j occ vcoul[igp(j)] sum_ig ssx*Mo
// with P = delw*I_eps (noflagOCC_solver, :139-146) and sch, ssx from
// ssxt_scht_solver (:16-63); sch = P in branch 1 and 0 otherwise.
#pragma once

#include <type_traits>

#include "gpp_internal.hpp"

namespace gpp {
namespace fact {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kKS = 8;             // bands per LDS stage of the smaller FP64 Gram tiles (two MFMA k-steps)
constexpr int kWJ = 64, kWI = 32;  // FP64 Gram wave tile: 64 G' columns x 32 G-vectors
constexpr int kRedThreads = 256;   // pole-reduce and final-reduce blocks

// FP64 Gram block tile made of WJN x WIN waves; KS bands per LDS stage,
// staged through registers in NH pieces.  The 128 x 128 tile stages 16 bands
// (four k-steps per barrier) in two pieces of 8, each stored to the other
// buffer after its k-steps, so its staging registers fit two waves per SIMD;
// the smaller tiles stage 8 in one piece (more would spill).
template <int WJN, int WIN> struct GramCfg {
    static constexpr int TJ = kWJ * WJN, TI = kWI * WIN, NT = 64 * WJN * WIN;
    static constexpr int KS = WJN * WIN == 8 ? 16 : kKS;
    static constexpr int NH = WJN * WIN == 8 ? 2 : 1;
};

// Stage 2: one TJ x TI tile of Mu = conj(Bg)^T * aqsn over bands [k_begin, k_end)
// on v_mfma_f64_16x16x4f64.  Lane layout (measured on gfx950): A[l%16][l/16],
// B[l/16][l%16], result r of lane l = C[4r + l/16][l%16]; MFMA rows are G'
// columns j and MFMA columns are G-vectors ig, so each result row is
// contiguous in ig.  sB/sA are the block's double-buffered LDS stages.
// Three real products per complex product (3M): with b = Bg, a = aqsn,
//   P1 = sum br ar,  P2 = sum bi ai,  P3 = sum (br - bi)(ar + ai),
//   Re Mu = P1 + P2,  Im Mu = (P3 - P1) + P2  (= sum br ai - bi ar),
// three accumulators fed by one MFMA each per k-step instead of two fed by
// two.  Rounding: |P3| <= 2 sum |b||a|, so the error stays of order
// 1e-16 sum |b||a| (normwise), as with four products.
//
// TRANS swaps the MFMA operands: rows are then G-vectors and columns G'
// columns, result r of lane l of sub-tile (s, t) is element
// (ig = wi + 16t + 4r + l/16, j = wj + 16s + l%16) -- each lane holds four
// G' columns and eight G-vectors of each, the layout the fused pole
// epilogue sums over G-vectors in.  Every element is the same sum of the
// same products in either orientation.
template <int WJN, int WIN, bool TRANS = false>
__device__ __forceinline__ void gram_f64_acc(const cplx *__restrict__ bg, const cplx *__restrict__ aqsn, int ngpown,
                                             int ncouls, int k_begin, int k_end, int j0, int i0,
                                             cplx (*sB)[GramCfg<WJN, WIN>::KS][GramCfg<WJN, WIN>::TJ],
                                             cplx (*sA)[GramCfg<WJN, WIN>::KS][GramCfg<WJN, WIN>::TI], v4d (&acc1)[4][2],
                                             v4d (&acc2)[4][2], v4d (&acc3)[4][2])
{
    typedef GramCfg<WJN, WIN> C;
    constexpr int TJ = C::TJ, TI = C::TI, NT = C::NT, KS = C::KS, NH = C::NH;
    constexpr int PB = KS * TJ / NT / NH, PA = KS * TI / NT / NH;  // staged elements per thread and piece
    static_assert(PB * NT * NH == KS * TJ && PA * NT * NH == KS * TI, "stage must split evenly");
    static_assert((PB * NT) % TJ == 0 && (PA * NT) % TI == 0 && (KS / NH) % 4 == 0,
                  "a piece is whole band rows and whole k-steps");
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    const int wj = (wave / WIN) * kWJ;
    const int wi = (wave % WIN) * kWI;

#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int t = 0; t < 2; ++t) acc1[s][t] = acc2[s][t] = acc3[s][t] = v4d{0, 0, 0, 0};

    // piece h of a stage: band rows [h KS / NH, (h + 1) KS / NH)
    cplx stB[PB], stA[PA];
    auto load_piece = [&](int k0, int h) {
#pragma unroll
        for (int q = 0; q < PB; ++q) {
            const int e = threadIdx.x + (h * PB + q) * NT;
            const int kr = e / TJ, c = e % TJ, k = k0 + kr, j = j0 + c;
            stB[q] = (k < k_end && j < ngpown) ? bg[(size_t)k * ngpown + j] : czero();
        }
#pragma unroll
        for (int q = 0; q < PA; ++q) {
            const int e = threadIdx.x + (h * PA + q) * NT;
            const int kr = e / TI, c = e % TI, k = k0 + kr, ig = i0 + c;
            stA[q] = (k < k_end && ig < ncouls) ? aqsn[(size_t)k * ncouls + ig] : czero();
        }
    };
    auto store_piece = [&](int buf, int h) {
#pragma unroll
        for (int q = 0; q < PB; ++q) {
            const int e = threadIdx.x + (h * PB + q) * NT;
            sB[buf][e / TJ][e % TJ] = stB[q];
        }
#pragma unroll
        for (int q = 0; q < PA; ++q) {
            const int e = threadIdx.x + (h * PA + q) * NT;
            sA[buf][e / TI][e % TI] = stA[q];
        }
    };

    const int nst = (k_end - k_begin + KS - 1) / KS;
#pragma unroll
    for (int h = 0; h < NH; ++h) {
        load_piece(k_begin, h);
        store_piece(0, h);
    }
    __syncthreads();
    for (int st = 0; st < nst; ++st) {
        const int buf = st & 1;
#pragma unroll
        for (int h = 0; h < NH; ++h) {
        if (st + 1 < nst) load_piece(k_begin + (st + 1) * KS, h);
        __builtin_amdgcn_s_setprio(1);  // the MFMA section: the wave with ready operands issues first
        // scheduling hint: interleave the fragment reads from LDS with the MFMAs
        // (strategy 1; measured ~1 % faster than the default schedule and than strategy 0)
        __builtin_amdgcn_iglp_opt(1);
#pragma unroll
        for (int kk = h * (KS / NH); kk < (h + 1) * (KS / NH); kk += 4) {
            const int kr = kk + lane / 16;
            double are[2], aim[2], asum[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const cplx a = sA[buf][kr][wi + 16 * t + lane % 16];
                are[t] = a.re;
                aim[t] = a.im;
                asum[t] = a.re + a.im;
            }
            // one B sub-tile at a time keeps the fragments within the
            // register budget of two waves per SIMD
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const cplx b = sB[buf][kr][wj + 16 * s + lane % 16];
                const double bdiff = b.re - b.im;
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    if constexpr (TRANS) {
                        acc1[s][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(are[t], b.re, acc1[s][t], 0, 0, 0);
                        acc2[s][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(aim[t], b.im, acc2[s][t], 0, 0, 0);
                        acc3[s][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(asum[t], bdiff, acc3[s][t], 0, 0, 0);
                    } else {
                        acc1[s][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(b.re, are[t], acc1[s][t], 0, 0, 0);
                        acc2[s][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(b.im, aim[t], acc2[s][t], 0, 0, 0);
                        acc3[s][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(bdiff, asum[t], acc3[s][t], 0, 0, 0);
                    }
                }
            }
        }
        __builtin_amdgcn_s_setprio(0);
        if (st + 1 < nst) store_piece(buf ^ 1, h);
        }
        __syncthreads();
    }
}

// Stage 2 with Mu written to HBM ([ngpown][ncouls], rows contiguous in ig).
template <int WJN, int WIN>
__device__ __forceinline__ void gram_f64_tile(const cplx *__restrict__ bg, const cplx *__restrict__ aqsn, int ngpown,
                                              int ncouls, int k_begin, int k_end, int j0, int i0,
                                              cplx (*sB)[GramCfg<WJN, WIN>::KS][GramCfg<WJN, WIN>::TJ],
                                              cplx (*sA)[GramCfg<WJN, WIN>::KS][GramCfg<WJN, WIN>::TI],
                                              cplx *__restrict__ mout)
{
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    const int wj = (wave / WIN) * kWJ;
    const int wi = (wave % WIN) * kWI;
    v4d acc1[4][2], acc2[4][2], acc3[4][2];  // P1, P2, P3
    gram_f64_acc<WJN, WIN>(bg, aqsn, ngpown, ncouls, k_begin, k_end, j0, i0, sB, sA, acc1, acc2, acc3);
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int j = j0 + wj + 16 * s + 4 * r + lane / 16;
                const int ig = i0 + wi + 16 * t + lane % 16;
                if (j < ngpown && ig < ncouls) {
                    // non-temporal: Mu is written once and read back once, by
                    // the pole reduce (C1 reduce 0.486-0.502 -> 0.469 ms)
                    __builtin_nontemporal_store(acc1[s][t][r] + acc2[s][t][r], &mout[(size_t)j * ncouls + ig].re);
                    __builtin_nontemporal_store((acc3[s][t][r] - acc1[s][t][r]) + acc2[s][t][r],
                                                &mout[(size_t)j * ncouls + ig].im);
                }
            }
}

// 2^-40 band around a threshold inside which a decision made on an estimate
// is remade in the reference's arithmetic.
constexpr double kBand = 1.0 / 1099511627776.0;

// The occupied solver's branch tests (gppComplex.cpp:35, :43) on |delw|^2 for
// one element and frequency: branch1 = |wdiff|^2 > 0.25 and |delw|^2 <
// limitone, above2 = |delw|^2 > to1.  wdiffr = |wdiff|^2 rounded as the
// reference rounds it, rden its reciprocal (rcp_1nr), wt2 = |wt|^2.
// est = wt2 * rden is within 1e-14 of the reference's rounded |delw|^2,
// relative to it (each rounding of delw is bounded by |delw|, with or
// without cancellation), so outside a 2^-40 band around the two thresholds
// it decides as the reference does; inside it, below the lower one and for a
// NaN est (|wdiff|^2 = 0 or beyond the double range: rcp_1nr returns NaN),
// the reference's delw and |delw|^2 are recomputed in its operation order
// with the correctly rounded reciprocal (:30-32).  An est of +inf (the
// product overflowed) means a |delw|^2 of at least about 1.8e308, far above
// both thresholds, where the reference decides alike (branch 2).
__device__ __forceinline__ void occ_branch_tests(double wxf, cplx wt, double wdiffr, double rden, double wt2,
                                                 bool &branch1, bool &above2)
{
#pragma clang fp contract(off)
    constexpr double limitone = 1.0 / (kTo1 * 4.0);
    constexpr double kEstLo1 = limitone * (1 - kBand), kEstHi1 = limitone * (1 + kBand);
    constexpr double kEstHi2 = kTo1 * (1 + kBand);
    const double est = wt2 * rden;
    // below the to1 band (|delw|^2 < 1e-6, where |wdiff| exceeds |wt| a
    // thousandfold) is rare enough to take the exact path with the bands
    const bool sure = ((est > kEstHi2) & (est < kEstLo1)) | (est > kEstHi1);
    bool below1 = est < kEstLo1;
    above2 = est > kEstHi2;
    if (!sure) {
        const cplx wdiff = cdminus(wxf, wt);
        const cplx delw = cscale(cmul(wt, cconj(wdiff)), rcp_refined(wdiffr));
        const double delwr = cmul(delw, cconj(delw)).re;
        below1 = delwr < limitone;
        above2 = delwr > kTo1;
    }
    branch1 = (wdiffr > 0.25) && below1;
}

// One instance's inputs and workspace as seen by the reduce stages.
struct ReduceArgs {
    const cplx *mu;  // [ngpown][ncouls]
    const cplx *ieps, *wtilde, *aqsn;
    const cplx *bg;  // [nb][ngpown]
    const int32_t *idx;
    const double *vcoul, *wx;
    int nv, ncouls, ngpown, nfreq, nsplit, chunk;
    int fpb;  // frequencies per block (single-instance grid: z = ceil(nfreq / fpb))
    double occ;
    cplx *col_ach, *col_asx;  // [ngpown][nsplit][nfreq]
    // Optional (acht_n1_loc): the pole factor of frequency acht_iw, p2 =
    // delw*I_eps [ngpown][ncouls], and br1 [ngpown][ncouls] = 1 where the
    // occupied solver takes branch 1 at that frequency (sch = p2 there, 0
    // elsewhere); p2 null when not requested, br1 null when nv == 0.
    cplx *p2;
    uint8_t *br1;
    int acht_iw;
    int neg;  // some frequency is negative: launch the NEG variant (host-side dispatch only)
    // With p2 and reduce_sums_acht_occ: the occupied bands' acht_n1_loc
    // partials, [ngpown][nsplit][nv], summed in the pass (br1 not written).
    cplx *part_occ;
};

// Whether the p2-storing reduce sums the occupied bands' acht_n1_loc terms
// itself or leaves them to acht_occ_item (br1): two occupied bands, whose
// aqsn values and sums fit beside the pass (with 4 or 8 the kernels spill,
// at 8 and 16 frequencies by 45-297 registers).
__host__ __device__ inline bool reduce_sums_acht_occ(int nv, int neg) { return !neg && nv == 2; }

// Stage 3: pole terms of column j over G-vectors [split*chunk, +chunk), for
// frequencies [iw0, iw0+FC).  Per element and frequency: wdiff, |wdiff|^2,
// delw = wt*conj(wdiff)*rden and |delw|^2 in the reference's order
// (gppComplex.cpp:139-144, no contraction; rden is the refined reciprocal,
// equal to the IEEE quotient), so the occupied branch tests (:35, :43) see
// the reference's values.  P = delw*ie, the branch-2 ssx in its reduced form
// -(I_eps wt / 2) / (wx + wt) and the accumulations use FMA.  sch = P in branch 1 (:37) shares the
// unoccupied product.  Block-wide fixed-order sums; writes the column
// partials (0.5*sum)*vcoul and sum*(occ*vcoul).  scratch: 4 entries.
// NV > 0: the occupied band count is NV (== a.nv), known at compile time, so
// the occupied loads are unrolled and issued together and conj(bg) of the
// column is loaded once; NV == 0: any a.nv.  NEG: some frequency may be
// negative, where the ssx cutoff of :58-59 is live; without it every
// frequency must be >= 0 (the launcher checks on the host) and the kernel
// carries no cutoff code.  P2 with NV == 2 (and no NEG): the pass holding
// frequency acht_iw also sums sum_ig aqsn[n1,ig] p2[j,ig] [branch 1] per
// occupied band n1 (the same products in the same order as acht_occ_item,
// whose partials it writes to part_occ).
template <int FC, int NV = 0, bool P2 = false, int FCT = FC, bool NEG = false>
__device__ __forceinline__ void pole_reduce_item(const ReduceArgs &a, int j, int split, int iw_begin, int iw_end,
                                                 cplx *scratch)
{
#pragma clang fp contract(off)
    const int ncouls = a.ncouls;
    const int nv = NV > 0 ? NV : a.nv;
    constexpr bool kOccSum = P2 && NV == 2 && !NEG;  // reduce_sums_acht_occ

    const int ig_begin = split * a.chunk;
    const int ig_end = ig_begin + a.chunk < ncouls ? ig_begin + a.chunk : ncouls;
    // conj(Bg[n1][j]) of the occupied bands, loaded once for the column (the
    // optional p2/br1 stores in the loop would otherwise keep the compiler
    // from hoisting these loads out of it).
    // Wave-uniform (the column's), so kept in scalar registers.
    cplx bgc[NV > 0 ? NV : 1];
    if constexpr (NV > 0) {
#pragma unroll
        for (int n1 = 0; n1 < NV; ++n1) bgc[n1] = uniform(cconj(a.bg[(size_t)n1 * a.ngpown + j]));
    }
    // Frequencies in passes of FC, the last one FCT wide when at most FCT
    // remain; the passes after the first re-read this block's G-vector range,
    // recently fetched, from the caches.
    auto pass = [&](auto width, auto fold, int iw0) {
        constexpr int F = decltype(width)::value;
        constexpr bool FOLD = decltype(fold)::value;
        double wxv[F], c1[F];  // branch-1 ssx = Omega2 * wx^2 / wx^4 (:38-41) = Omega2 * c1
#pragma unroll
        for (int f = 0; f < F; ++f) {
            // wave-uniform, so kept in scalar registers
            wxv[f] = uniform((iw0 + f < iw_end) ? a.wx[iw0 + f] : 1.0);
            const double wx2 = wxv[f] * wxv[f];
            c1[f] = uniform(wx2 * (1.00 / (wx2 * wx2)));
        }
        cplx sach[F], sasx[F];
#pragma unroll
        for (int f = 0; f < F; ++f) sach[f] = sasx[f] = czero();
        cplx oacc[kOccSum ? NV : 1];  // occupied acht_n1_loc sums (kOccSum)
#pragma unroll
        for (int n1 = 0; n1 < (kOccSum ? NV : 1); ++n1) oacc[n1] = czero();

        bool any_neg = false;  // the ssx cutoff of :58-59 applies only at negative frequencies
        if constexpr (NEG) {
#pragma unroll
            for (int f = 0; f < F; ++f) any_neg = any_neg || wxv[f] < 0.00;
        }

        const size_t row = (size_t)j * ncouls;
        // The next element's wtilde, I_eps and Mu load while this one is
        // evaluated (two elements in flight per thread; a third did not
        // help).
        // Kernels of at most three frequencies make one pass: their three
        // streams are read once, without keeping them in the caches.
        auto ld = [](const cplx *p) -> cplx {
            if constexpr (FC <= 3)
                return cplx{__builtin_nontemporal_load(&p->re), __builtin_nontemporal_load(&p->im)};
            else
                return *p;
        };
        cplx wt_n = czero(), ie_n = czero(), mu_n = czero();
        if (ig_begin + (int)threadIdx.x < ig_end) {
            wt_n = ld(a.wtilde + row + ig_begin + threadIdx.x);
            ie_n = ld(a.ieps + row + ig_begin + threadIdx.x);
            mu_n = ld(a.mu + row + ig_begin + threadIdx.x);
        }
        for (int ig = ig_begin + threadIdx.x; ig < ig_end; ig += blockDim.x) {
            const cplx wt = wt_n, ie = ie_n, mu = mu_n;
            if (ig + (int)blockDim.x < ig_end) {
                wt_n = ld(a.wtilde + row + ig + blockDim.x);
                ie_n = ld(a.ieps + row + ig + blockDim.x);
                mu_n = ld(a.mu + row + ig + blockDim.x);
            }
            // P = delw I_eps = q conj(wdiff) / |wdiff|^2 with q = wt I_eps and
            // conj(wdiff) = (wx - wt.re, wt.im) =: (wdr, wt.im).  FOLD (problems
            // with 4 or more frequencies) folds Mu and Mo into the products
            // once per element -- u = q Mu (every element), uo = q Mo (branch
            // 1: sch = P for the occupied bands) and Omega2 Mo (branch 1's ssx
            // Mo = c1 Omega2 Mo) -- so each frequency adds u conj(wdiff) /
            // |wdiff|^2 with four FMAs (C4 reduce 18.2 -> 14.9 ms); with fewer
            // frequencies P is formed and Mu or Mu + Mo selected per
            // frequency, which is faster there (C1: 0.50 against 0.55 ms
            // folded).
            const cplx q = cmul_fast(wt, ie);
            // wdiff.im^2 rounded as the reference's |wdiff|^2 (:27, :33) rounds it
            const double wti2 = wt.im * wt.im;
            // q.im wt.im and q.re wt.im: kept per element when P is formed every
            // frequency, recomputed on the rare paths otherwise
            const double q_im_wti_e = FOLD ? 0.0 : q.im * wt.im, q_re_wti_e = FOLD ? 0.0 : q.re * wt.im;
            cplx mo = czero(), om2 = czero(), mum = mu, u0 = czero(), uo = czero(), w1 = czero();
            double u0_im_wti = 0.0, u0_re_wti = 0.0, uo_im_wti = 0.0, uo_re_wti = 0.0, ie_cut = 0.0, wt2 = 0.0;
            if constexpr (FOLD) {
                u0 = cmul_fast(q, mu);
                u0_im_wti = u0.im * wt.im;
                u0_re_wti = u0.re * wt.im;
            }
            cplx av[NV > 0 ? NV : 1];  // the occupied aqsn of this G-vector, loads issued together
            if (nv > 0) {
                // Mo over the occupied bands, in band order: folded passes
                // take four FMAs per band (C3 reduce 7.54 -> 7.33 ms, C4 14.39
                // -> 14.14 ms), the three-frequency kernel a product and a sum
                // (0.487 against 0.501 ms at 512/2/32768/1638)
                if constexpr (NV > 0) {
#pragma unroll
                    for (int n1 = 0; n1 < NV; ++n1) av[n1] = a.aqsn[(size_t)n1 * ncouls + ig];
                }
                if constexpr (NV > 0 && FOLD) {
                    mo = cmul_fast(bgc[0], av[0]);
#pragma unroll
                    for (int n1 = 1; n1 < NV; ++n1) mo = cfma(bgc[n1], av[n1], mo);
                } else if constexpr (NV > 0) {
#pragma unroll
                    for (int n1 = 0; n1 < NV; ++n1) mo = cadd(mo, cmul_fast(bgc[n1], av[n1]));
                } else if constexpr (FOLD) {
                    for (int n1 = 0; n1 < nv; ++n1)
                        mo = cfma(cconj(a.bg[(size_t)n1 * a.ngpown + j]), a.aqsn[(size_t)n1 * ncouls + ig], mo);
                } else {
                    for (int n1 = 0; n1 < nv; ++n1)
                        mo = cadd(mo, cmul_fast(cconj(a.bg[(size_t)n1 * a.ngpown + j]),
                                                a.aqsn[(size_t)n1 * ncouls + ig]));
                }
                om2 = cmul(csquare(wt), ie);  // Omega2 = wtilde^2 * I_eps, flagOCC_solver :117-118
                if constexpr (FOLD) {
                    uo = cmul_fast(q, mo);
                    uo_im_wti = uo.im * wt.im;
                    uo_re_wti = uo.re * wt.im;
                    w1 = cmul_fast(om2, mo);
                } else {
                    mum = cadd(mu, mo);
                }
                wt2 = __builtin_fma(wt.re, wt.re, wti2);
                if (NEG && any_neg) {
                    // A real (uniform) branch: the empty asm keeps the compiler
                    // from evaluating the square root for every element.
                    asm volatile("" ::: "memory");
                    ie_cut = cabs_(ie) * kSexcut;
                }
            }
#pragma unroll
            for (int f = 0; f < F; ++f) {
                const double wdr = wxv[f] - wt.re;        // wdiff = (wdr, -wt.im), :27
                const double wdiffr = wdr * wdr + wti2;   // |wdiff|^2 exactly as the reference rounds it
                const double rden = rcp_1nr(wdiffr);
                if constexpr (FOLD) {  // P Mu
                    sach[f].re = __builtin_fma(__builtin_fma(u0.re, wdr, -u0_im_wti), rden, sach[f].re);
                    sach[f].im = __builtin_fma(__builtin_fma(u0.im, wdr, u0_re_wti), rden, sach[f].im);
                }
                bool branch1 = false, above2 = false;
                if (nv > 0) occ_branch_tests(wxv[f], wt, wdiffr, rden, wt2, branch1, above2);
                // Branch 2 (:46-50) with cden = 4 wt^2 (0.5 + delw), Omega2 =
                // wt^2 I_eps and delw = wt / (wx - wt):
                //   -Omega2 conj(cden) delw / |cden|^2 = -I_eps delw / (2 + 4 delw)
                //                                     = -(q / 2) / (wx + wt),
                // a pole at -wt, far from cancellation for wx > 0, with
                // wx + wt = (s, wt.im).
                const bool neg_f = NEG && wxv[f] < 0.00;
                if (!FOLD || neg_f) {
                    const double q_im_wti = FOLD ? q.im * wt.im : q_im_wti_e;
                    const double q_re_wti = FOLD ? q.re * wt.im : q_re_wti_e;
                    const cplx p = {__builtin_fma(q.re, wdr, -q_im_wti) * rden,
                                    __builtin_fma(q.im, wdr, q_re_wti) * rden};
                    if (nv > 0) {
                        cplx ssx = czero();
                        if (branch1) {
                            ssx = cscale(om2, c1[f]);
                        } else if (above2) {
                            const double s = wxv[f] + wt.re;
                            const double rr = 0.5 * rcp_1nr(__builtin_fma(s, s, wti2));
                            ssx = cplx{-__builtin_fma(q.re, s, q_im_wti) * rr, -__builtin_fma(q.im, s, -q_re_wti) * rr};
                        }
                        // The ssx cutoff of :58-59 (only at negative
                        // frequencies), decided on the reference's |ssx|.
                        // Branch 1: the reference's ssx is ((Omega2 wx^2)
                        // (1/wx^4)) (:38-41) with Omega2 already in its
                        // operation order.  Branch 2: this ssx is within about
                        // 2^-45 of the reference's, relative (rcp_1nr, the
                        // reduced form), so outside a 2^-40 band around the
                        // cutoff it decides as the reference does; inside it
                        // the reference's ssx is recomputed and decides.
                        if (neg_f) {
                            bool cut;
                            if (branch1) {
                                const double wx2 = wxv[f] * wxv[f];
                                cut = cabs_(cscale2(om2, wx2, 1.00 / (wx2 * wx2))) > ie_cut;
                            } else {
                                const double ax = cabs_(ssx);
                                cut = ax > ie_cut;
                                if (__builtin_fabs(ax - ie_cut) <= kBand * ie_cut)
                                    cut = occ_branch2_abs_ssx(wxv[f], wt, ie) > ie_cut;
                            }
                            if (cut) ssx = czero();
                        }
                        sasx[f] = cfma(ssx, mo, sasx[f]);
                    }
                    if constexpr (FOLD) {
                        if (branch1) sach[f] = cfma(p, mo, sach[f]);  // P Mu is in already
                    } else {
                        sach[f] = cfma(p, branch1 ? mum : mu, sach[f]);
                    }
                } else if (branch1) {  // P Mo and c1 Omega2 Mo
                    sach[f].re = __builtin_fma(__builtin_fma(uo.re, wdr, -uo_im_wti), rden, sach[f].re);
                    sach[f].im = __builtin_fma(__builtin_fma(uo.im, wdr, uo_re_wti), rden, sach[f].im);
                    sasx[f].re = __builtin_fma(c1[f], w1.re, sasx[f].re);
                    sasx[f].im = __builtin_fma(c1[f], w1.im, sasx[f].im);
                } else if (above2) {  // -(uo / 2) conj(wx + wt) / |wx + wt|^2
                    const double s = wxv[f] + wt.re;
                    const double rr = -0.5 * rcp_1nr(__builtin_fma(s, s, wti2));
                    sasx[f].re = __builtin_fma(rr, __builtin_fma(uo.re, s, uo_im_wti), sasx[f].re);
                    sasx[f].im = __builtin_fma(rr, __builtin_fma(uo.im, s, -uo_re_wti), sasx[f].im);
                }
                if (P2 && iw0 + f == a.acht_iw && iw0 + f < iw_end) {
                    const cplx pv = {__builtin_fma(q.re, wdr, -(FOLD ? q.im * wt.im : q_im_wti_e)) * rden,
                                     __builtin_fma(q.im, wdr, FOLD ? q.re * wt.im : q_re_wti_e) * rden};
                    __builtin_nontemporal_store(pv.re, &a.p2[row + ig].re);  // streamed, read once by the acht product
                    __builtin_nontemporal_store(pv.im, &a.p2[row + ig].im);
                    if constexpr (kOccSum) {
                        if (branch1) {  // sch = P (:37)
#pragma unroll
                            for (int n1 = 0; n1 < (kOccSum ? NV : 1); ++n1) oacc[n1] = cfma(av[n1], pv, oacc[n1]);
                        }
                    } else if (a.br1) {
                        a.br1[row + ig] = branch1 ? 1 : 0;
                    }
                }
            }
        }
        const double vc = a.vcoul[igp_of(a.idx, j, ncouls)];
#pragma unroll
        for (int f = 0; f < F; ++f) {
            const cplx s = block_sum(sach[f], scratch);
            const cplx x = block_sum(sasx[f], scratch);
            if (threadIdx.x == 0 && iw0 + f < iw_end) {
                // (0.5 * sum) * vcoul and sum * (occ * vcoul), gppComplex.cpp:361-363, :379
                const size_t o = ((size_t)j * a.nsplit + split) * a.nfreq + iw0 + f;
                a.col_ach[o] = cscale(cscale(s, 0.5), vc);
                a.col_asx[o] = cscale(x, a.occ * vc);
            }
        }
        if constexpr (kOccSum) {
            if (a.acht_iw >= iw0 && a.acht_iw < iw0 + F && a.acht_iw < iw_end) {  // uniform
                // 0.5 vcoul[igp(j)] conj(aqsm[n1,igp(j)]) sum, gppComplex.cpp:381
                const double w = 0.5 * vc;
#pragma unroll
                for (int n1 = 0; n1 < (kOccSum ? NV : 1); ++n1) {
                    const cplx s = block_sum(oacc[n1], scratch);
                    if (threadIdx.x == 0)
                        a.part_occ[((size_t)j * a.nsplit + split) * NV + n1] = cscale(cmul_fast(bgc[n1], s), w);
                }
            }
        }
    };
    // The choice depends on the problem's frequency count only, so every
    // grouping of its frequencies into passes (and blocks) sums alike.
    using fold_off = std::integral_constant<bool, false>;
    using fold_on = std::integral_constant<bool, true>;
    int iw0 = iw_begin;
    if constexpr (FC <= 3) {
        for (; iw_end - iw0 > FCT; iw0 += FC) pass(std::integral_constant<int, FC>(), fold_off(), iw0);
        if (iw0 < iw_end) pass(std::integral_constant<int, FCT>(), fold_off(), iw0);
    } else if (a.nfreq < 4) {  // at most 3 frequencies: one pass FCT (>= 4) wide
        if (iw0 < iw_end) pass(std::integral_constant<int, FCT>(), fold_off(), iw0);
    } else if constexpr (FC >= 16) {
        // passes of 16 while more than 12 remain, then at most one of 8 and
        // one of 4, so no pass runs more than 3 idle frequencies beyond the
        // first (24: 16 + 8 instead of 16 + 16)
        for (; iw_end - iw0 > 12; iw0 += 16) pass(std::integral_constant<int, 16>(), fold_on(), iw0);
        if (iw_end - iw0 > 4) {
            pass(std::integral_constant<int, 8>(), fold_on(), iw0);
            iw0 += 8;
        }
        if (iw0 < iw_end) pass(std::integral_constant<int, 4>(), fold_on(), iw0);
    } else if constexpr (FC == 8) {
        // passes of 8 while more than 8 remain, then one of 8, 6, 4 or 2: at
        // most one idle frequency (13: 8 + 6; C4 reduce 11.5 -> 11.1 ms
        // against a last pass of 4 or 8)
        for (; iw_end - iw0 > 8; iw0 += 8) pass(std::integral_constant<int, 8>(), fold_on(), iw0);
        const int rem = iw_end - iw0;
        if (rem > 6)
            pass(std::integral_constant<int, 8>(), fold_on(), iw0);
        else if (rem > 4)
            pass(std::integral_constant<int, 6>(), fold_on(), iw0);
        else if (rem > 2)
            pass(std::integral_constant<int, 4>(), fold_on(), iw0);
        else if (rem > 0)
            pass(std::integral_constant<int, 2>(), fold_on(), iw0);
    } else {
        for (; iw_end - iw0 > FCT; iw0 += FC) pass(std::integral_constant<int, FC>(), fold_on(), iw0);
        if (iw0 < iw_end) pass(std::integral_constant<int, FCT>(), fold_on(), iw0);
    }
}

// One element's pole terms at F frequencies, evaluated as pole_reduce_item
// does with fewer than four frequencies and none negative: P = delw I_eps
// formed per frequency, sch = P Mu, or P (Mu + Mo) in branch 1; ssx from the
// branch (branch 1: c1 Omega2; branch 2: the reduced form -(q / 2) /
// (wx + wt)) times Mo.  occ: the problem has occupied bands.
template <int F>
__device__ __forceinline__ void pole_element_plain(const double (&wxv)[F], const double (&c1)[F], bool occ, cplx wt,
                                                   cplx ie, cplx mu, cplx mo, cplx (&sach)[F], cplx (&sasx)[F])
{
#pragma clang fp contract(off)
    const cplx q = cmul_fast(wt, ie);
    const double wti2 = wt.im * wt.im;  // rounded as the reference's |wdiff|^2 rounds it
    const double q_im_wti = q.im * wt.im, q_re_wti = q.re * wt.im;
    cplx om2 = czero(), mum = mu;
    double wt2 = 0.0;
    if (occ) {
        om2 = cmul(csquare(wt), ie);  // Omega2 = wtilde^2 * I_eps, flagOCC_solver :117-118
        mum = cadd(mu, mo);
        wt2 = __builtin_fma(wt.re, wt.re, wti2);
    }
#pragma unroll
    for (int f = 0; f < F; ++f) {
        const double wdr = wxv[f] - wt.re;       // wdiff = (wdr, -wt.im), :27
        const double wdiffr = wdr * wdr + wti2;  // |wdiff|^2 exactly as the reference rounds it
        const double rden = rcp_1nr(wdiffr);
        bool branch1 = false, above2 = false;
        if (occ) occ_branch_tests(wxv[f], wt, wdiffr, rden, wt2, branch1, above2);
        const cplx p = {__builtin_fma(q.re, wdr, -q_im_wti) * rden, __builtin_fma(q.im, wdr, q_re_wti) * rden};
        if (occ) {
            cplx ssx = czero();
            if (branch1) {
                ssx = cscale(om2, c1[f]);
            } else if (above2) {
                const double sr = wxv[f] + wt.re;
                const double rr = 0.5 * rcp_1nr(__builtin_fma(sr, sr, wti2));
                ssx = cplx{-__builtin_fma(q.re, sr, q_im_wti) * rr, -__builtin_fma(q.im, sr, -q_re_wti) * rr};
            }
            sasx[f] = cfma(ssx, mo, sasx[f]);
        }
        sach[f] = cfma(p, branch1 ? mum : mu, sach[f]);
    }
}

// Inputs and outputs of the Gram product with the pole reduce fused into its
// epilogue (at most 3 frequencies, none negative, acht_n1_loc not requested).
struct PoleTileArgs {
    const cplx *aqsn, *ieps, *wtilde;
    const cplx *bg;  // [nb][ngpown]
    const double *wx;
    int nb, nv, ncouls, ngpown, nsplit;  // nsplit = G-vector tiles
    cplx *col_ach, *col_asx;             // [ngpown][nsplit][F], unscaled tile sums
};

// LDS of the fused Gram + pole reduce block (128 x 128 tile, 8 waves), in
// complex entries: the Gram's double-buffered band stages (sB, sA: 4096
// entries) during the product; afterwards, for the epilogue rounds, wtilde,
// I_eps and Mu of 32 columns x 64 G-vectors (rows padded to 65 entries, so
// the transposed reads of 8 lanes fall in 8 different 16-byte bank groups),
// conj(Bg) and aqsn of the tile's occupied bands (at most kFuseNv), and the
// per-wave column sums of one column round.
constexpr int kFuseNv = 4;
template <int F> struct FuseLds {
    typedef GramCfg<2, 4> C;
    static constexpr int kCols = 32, kRows = 64;   // columns x G-vectors per round
    static constexpr int kPitch = kRows + 1;       // entries per staged row
    static constexpr int kWt = 0, kIe = kCols * kPitch, kMu = 2 * kCols * kPitch;
    static constexpr int kBg = 3 * kCols * kPitch;     // [kFuseNv][TJ]
    static constexpr int kAq = kBg + kFuseNv * C::TJ;  // [kFuseNv][TI]
    static constexpr int kRed = kAq + kFuseNv * C::TI;  // [WIN = 4][kCols][2F]
    static constexpr int kTotal = kRed + 4 * kCols * 2 * F;
    static constexpr int kGram = 2 * C::KS * (C::TJ + C::TI);  // the Gram's stage buffers
    static constexpr int kLds = kTotal > kGram ? kTotal : kGram;
};

// Epilogue of one 128 x 128 tile (G-vector tile `itile`) after
// gram_f64_acc<TRANS>: the accumulators are folded into Mu at once (128
// registers instead of 192), and the elements are evaluated in eight rounds
// (s, t): 32 columns (sub-tile s of both 64-column halves) x 64 G-vectors
// (sub-tile t of the four 32-G-vector quarters).  In a round all 512 threads
// stage its wtilde and I_eps rows into LDS with coalesced loads (the tile's
// occupied conj(Bg) and aqsn with the first), each wave writes its Mu of the
// round beside them, and each lane evaluates its four elements of the round
// in a loop -- Mo over the occupied bands (four FMAs per band), the pole
// terms at the F frequencies (pole_element_plain) -- summing per column over
// both rounds of a column sub-tile; then the four lane groups are summed by
// xor butterfly (every lane of a column ends with the same bits) and the
// four waves of a column range in wave order: one sum per (column, tile,
// frequency), written unscaled.
template <int F>
__device__ __forceinline__ void gram_pole_epilogue(const PoleTileArgs &a, int j0, int i0, int itile,
                                                   const v4d (&acc1)[4][2], const v4d (&acc2)[4][2],
                                                   const v4d (&acc3)[4][2], cplx *lds)
{
#pragma clang fp contract(off)
    typedef FuseLds<F> L;
    typedef GramCfg<2, 4> C;
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    const int wjn = wave / 4, wv = wave % 4;  // 64-column half, 32-G-vector quarter
    const int ncouls = a.ncouls, ngpown = a.ngpown, nv = a.nv;

    cplx mu[4][2][4];
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                mu[s][t][r] = cplx{acc1[s][t][r] + acc2[s][t][r], (acc3[s][t][r] - acc1[s][t][r]) + acc2[s][t][r]};

    double wxv[F], c1[F];  // branch-1 ssx = Omega2 * wx^2 / wx^4 (:38-41) = Omega2 * c1
#pragma unroll
    for (int f = 0; f < F; ++f) {
        wxv[f] = a.wx[f];
        const double wx2 = wxv[f] * wxv[f];
        c1[f] = wx2 * (1.00 / (wx2 * wx2));
    }
    // tile-local column of staged row c and G-vector of staged entry q in round (s, t)
    auto col_of = [](int c, int s) { return (c >> 4) * kWJ + 16 * s + (c & 15); };
    auto gv_of = [](int q, int t) { return (q >> 4) * kWI + 16 * t + (q & 15); };
    constexpr int PER = L::kCols * L::kRows / C::NT;  // staged entries per thread and array (4)
    const int jc = wjn * 16 + lane % 16;              // this lane's staged row

#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int jl = col_of(jc, s);
        cplx sach[F], sasx[F];
#pragma unroll
        for (int f = 0; f < F; ++f) sach[f] = sasx[f] = czero();
#pragma unroll
        for (int t = 0; t < 2; ++t) {
#pragma unroll
            for (int arr = 0; arr < 2; ++arr) {
                const cplx *src = arr ? a.ieps : a.wtilde;
                cplx v[PER];
#pragma unroll
                for (int k = 0; k < PER; ++k) {
                    const int e = threadIdx.x + k * C::NT;
                    const int j = j0 + col_of(e / L::kRows, s), ig = i0 + gv_of(e % L::kRows, t);
                    v[k] = (j < ngpown && ig < ncouls) ? src[(size_t)j * ncouls + ig] : czero();
                }
#pragma unroll
                for (int k = 0; k < PER; ++k) {
                    const int e = threadIdx.x + k * C::NT;
                    lds[(arr ? L::kIe : L::kWt) + (e / L::kRows) * L::kPitch + e % L::kRows] = v[k];
                }
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) lds[L::kMu + jc * L::kPitch + wv * 16 + 4 * r + lane / 16] = mu[s][t][r];
            if (s == 0 && t == 0) {
                for (int e = threadIdx.x; e < nv * C::TJ; e += C::NT) {
                    const int n = e / C::TJ, c = e % C::TJ;
                    lds[L::kBg + e] = j0 + c < ngpown ? cconj(a.bg[(size_t)n * ngpown + j0 + c]) : czero();
                    lds[L::kAq + e] = i0 + c < ncouls ? a.aqsn[(size_t)n * ncouls + i0 + c] : czero();
                }
            }
            __syncthreads();
            if (j0 + jl < ngpown) {
#pragma unroll 1
                for (int r = 0; r < 4; ++r) {
                    const int q = wv * 16 + 4 * r + lane / 16;  // staged entry
                    const int igl = gv_of(q, t);
                    if (i0 + igl < ncouls) {
                        cplx mo = czero();  // occupied bands in band order
                        for (int n1 = 0; n1 < nv; ++n1)
                            mo = cfma(lds[L::kBg + n1 * C::TJ + jl], lds[L::kAq + n1 * C::TI + igl], mo);
                        const int o = jc * L::kPitch + q;
                        pole_element_plain<F>(wxv, c1, nv > 0, lds[L::kWt + o], lds[L::kIe + o], lds[L::kMu + o], mo,
                                              sach, sasx);
                    }
                }
            }
            // every wave's reads of the round's rows are done before the next
            // round stages over them
            __syncthreads();
        }
#pragma unroll
        for (int f = 0; f < F; ++f) {
#pragma unroll
            for (int m = 16; m < kWave; m <<= 1) {
                sach[f].re += __shfl_xor(sach[f].re, m, kWave);
                sach[f].im += __shfl_xor(sach[f].im, m, kWave);
                sasx[f].re += __shfl_xor(sasx[f].re, m, kWave);
                sasx[f].im += __shfl_xor(sasx[f].im, m, kWave);
            }
        }
        if (lane < 16) {
#pragma unroll
            for (int f = 0; f < F; ++f) {
                lds[L::kRed + (wv * L::kCols + jc) * 2 * F + f] = sach[f];
                lds[L::kRed + (wv * L::kCols + jc) * 2 * F + F + f] = sasx[f];
            }
        }
        // the next column round writes its sums only after two more barriers
        __syncthreads();
        for (int e = threadIdx.x; e < L::kCols * 2 * F; e += C::NT) {
            const int c = e / (2 * F), k = e % (2 * F);
            const int j = j0 + col_of(c, s);
            if (j >= ngpown) continue;
            cplx v = lds[L::kRed + c * 2 * F + k];
#pragma unroll
            for (int w = 1; w < 4; ++w) v = cadd(v, lds[L::kRed + (w * L::kCols + c) * 2 * F + k]);
            const size_t o = ((size_t)j * a.nsplit + itile) * F + (k % F);
            if (k < F)
                a.col_ach[o] = v;
            else
                a.col_asx[o] = v;
        }
    }
}

// Column sums of the fused epilogue's tile sums: one wave per column j, the
// tiles summed in a fixed order (strided per lane, then xor butterfly), then
// scaled as the reduce scales them, (0.5 * sum) * vcoul and sum * (occ *
// vcoul) (gppComplex.cpp:361-363, :379); out [ngpown][nfreq].
__device__ __forceinline__ void pole_cols_wave(const cplx *col_ach, const cplx *col_asx, int nsplit, int nfreq,
                                               const int32_t *idx, const double *vcoul, int ncouls, double occ,
                                               int j, cplx *out_ach, cplx *out_asx)
{
    const int lane = threadIdx.x % kWave;
    const double vc = vcoul[igp_of(idx, j, ncouls)];
    for (int f = 0; f < nfreq; ++f) {
        cplx s = czero(), x = czero();
        for (int it = lane; it < nsplit; it += kWave) {
            s = cadd(s, col_ach[((size_t)j * nsplit + it) * nfreq + f]);
            x = cadd(x, col_asx[((size_t)j * nsplit + it) * nfreq + f]);
        }
        s = wave_sum(s);
        x = wave_sum(x);
        if (lane == 0) {
            out_ach[(size_t)j * nfreq + f] = cscale(cscale(s, 0.5), vc);
            out_asx[(size_t)j * nfreq + f] = cscale(x, occ * vc);
        }
    }
}

// acht_n1_loc (gppComplex.cpp:381, sch_array[2] of each band and column):
//   acht[n1] = sum_j 0.5 vcoul[igp(j)] conj(aqsm[n1,igp(j)]) sum_ig aqsn[n1,ig] R[j,ig]
// with R = p2 (P = delw*I_eps at that frequency) for the unoccupied bands and
// R = p2 masked by br1 (sch of ssxt_scht_solver, :37) for the occupied ones
// (acht_occ_item below: few bands, a streaming pass).  For the unoccupied
// bands Y = R * aqsn^T is a second Gram-type product (rows j, columns n1, K = G);
// one block computes a TJ x TN tile of Y over G-vectors [k_begin, k_end) on
// v_mfma_f64_16x16x4f64 with three real products per complex product (3M:
// P1 = sum rr ar, P2 = sum ri ai, P3 = sum (rr + ri)(ar + ai), Re Y = P1 - P2,
// Im Y = (P3 - P1) - P2) and reduces it at once to the column sums
// sum_j w_j conj(bg[n1][j]) Y[j][n1], written to part[n1] (one entry per
// column of the tile).  Operand tiles are staged K-fastest in 128-byte rows
// of 8 complex, element (row, k) at slot k ^ (row & 7): the stores (8 lanes
// per row) and the MFMA reads (16 rows per 16 lanes) are then free of LDS
// bank conflicts on gfx950 (measured with SQ_LDS_BANK_CONFLICT; a padded
// pitch of 9 complex is not).  A stage is two such pieces of 8 G-vectors
// (one barrier per 16), each loaded through registers during the previous
// piece's MFMAs and stored to the other buffer right after them.
constexpr int kAK = 8;          // G-vectors per LDS piece of the acht product
constexpr int kAKS = 2 * kAK;   // per stage

template <int WJN, int WIN> struct AchtCfg {
    static constexpr int TJ = kWJ * WJN, TN = kWI * WIN, NT = 64 * WJN * WIN;
};

template <int WJN, int WIN>
__device__ __forceinline__ void acht_f64_tile(const cplx *__restrict__ p2,
                                              const cplx *__restrict__ aqsn, const cplx *__restrict__ bg,
                                              const int32_t *__restrict__ idx, const double *__restrict__ vcoul,
                                              int ngpown, int ncouls, int n1_end, int k_begin, int k_end, int j0,
                                              int c0, cplx (*sR)[2][AchtCfg<WJN, WIN>::TJ][kAK],
                                              cplx (*sA)[2][AchtCfg<WJN, WIN>::TN][kAK],
                                              cplx (*red)[AchtCfg<WJN, WIN>::TN], cplx *__restrict__ part)
{
    typedef AchtCfg<WJN, WIN> C;
    constexpr int TJ = C::TJ, TN = C::TN, NT = C::NT;
    constexpr int PR = kAK * TJ / NT, PA = kAK * TN / NT;
    static_assert(PR * NT == kAK * TJ && PA * NT == kAK * TN, "stage must split evenly");
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    const int wjn = wave / WIN;
    const int wj = wjn * kWJ;
    const int wi = (wave % WIN) * kWI;

    v4d acc1[4][2], acc2[4][2], acc3[4][2];  // P1, P2, P3
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int t = 0; t < 2; ++t) acc1[s][t] = acc2[s][t] = acc3[s][t] = v4d{0, 0, 0, 0};

    cplx stR[PR], stA[PA];
    auto load_piece = [&](int k0) {
#pragma unroll
        for (int q = 0; q < PR; ++q) {
            const int e = threadIdx.x + q * NT;
            const int r = e / kAK, k = k0 + e % kAK, j = j0 + r;
            stR[q] = (k < k_end && j < ngpown) ? p2[(size_t)j * ncouls + k] : czero();
        }
#pragma unroll
        for (int q = 0; q < PA; ++q) {
            const int e = threadIdx.x + q * NT;
            const int r = e / kAK, k = k0 + e % kAK, n1 = c0 + r;
            stA[q] = (k < k_end && n1 < n1_end) ? aqsn[(size_t)n1 * ncouls + k] : czero();
        }
    };
    auto store_piece = [&](int buf, int h) {
#pragma unroll
        for (int q = 0; q < PR; ++q) {
            const int e = threadIdx.x + q * NT;
            const int r = e / kAK;
            sR[buf][h][r][(e % kAK) ^ (r & 7)] = stR[q];
        }
#pragma unroll
        for (int q = 0; q < PA; ++q) {
            const int e = threadIdx.x + q * NT;
            const int r = e / kAK;
            sA[buf][h][r][(e % kAK) ^ (r & 7)] = stA[q];
        }
    };

    const int nst = (k_end - k_begin + kAKS - 1) / kAKS;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        load_piece(k_begin + h * kAK);
        store_piece(0, h);
    }
    __syncthreads();
    for (int st = 0; st < nst; ++st) {
        const int buf = st & 1;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
        if (st + 1 < nst) load_piece(k_begin + (st + 1) * kAKS + h * kAK);
        // as in the Gram product: MFMA section first, fragment reads
        // interleaved with the MFMAs (C1 2.823 -> 2.799 ms, C3 33.96 -> 33.66)
        __builtin_amdgcn_s_setprio(1);
        __builtin_amdgcn_iglp_opt(1);
#pragma unroll
        for (int kk = 0; kk < kAK; kk += 4) {
            const int kr = (kk + lane / 16) ^ (lane & 7);  // every row read below is = lane (mod 8)
            double are[2], aim[2], asum[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const cplx a = sA[buf][h][wi + 16 * t + lane % 16][kr];
                are[t] = a.re;
                aim[t] = a.im;
                asum[t] = a.re + a.im;
            }
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const cplx r = sR[buf][h][wj + 16 * s + lane % 16][kr];
                const double rsum = r.re + r.im;
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    acc1[s][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(r.re, are[t], acc1[s][t], 0, 0, 0);
                    acc2[s][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(r.im, aim[t], acc2[s][t], 0, 0, 0);
                    acc3[s][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(rsum, asum[t], acc3[s][t], 0, 0, 0);
                }
            }
        }
        __builtin_amdgcn_s_setprio(0);
        if (st + 1 < nst) store_piece(buf ^ 1, h);
        }
        __syncthreads();
    }

    // Epilogue: result r of lane l is Y[j][n1] with j = 4r + l/16 + 16s + wj,
    // n1 = l%16 + 16t + wi.  Weight by 0.5 vcoul[igp(j)] conj(bg[n1][j]), sum
    // the wave's 64 rows (fixed order: s, r, then the xor butterfly over
    // l/16), then the WJN waves of one column range in wave order.
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int col = wi + 16 * t + lane % 16;
        const int n1 = c0 + col;
        cplx sum = czero();
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int j = j0 + wj + 16 * s + 4 * r + lane / 16;
                if (j < ngpown && n1 < n1_end) {
                    const double w = 0.5 * vcoul[igp_of(idx, j, ncouls)];
                    const cplx b = bg[(size_t)n1 * ngpown + j];
                    const cplx y = cplx{acc1[s][t][r] - acc2[s][t][r], (acc3[s][t][r] - acc1[s][t][r]) - acc2[s][t][r]};
                    sum = cadd(sum, cscale(cmul_fast(cconj(b), y), w));
                }
            }
        sum.re += __shfl_xor(sum.re, 16);
        sum.im += __shfl_xor(sum.im, 16);
        sum.re += __shfl_xor(sum.re, 32);
        sum.im += __shfl_xor(sum.im, 32);
        if (lane < 16) red[wjn][col] = sum;
    }
    __syncthreads();
    for (int col = threadIdx.x; col < TN; col += NT) {
        cplx s = czero();
#pragma unroll
        for (int w = 0; w < WJN; ++w) s = cadd(s, red[w][col]);
        if (c0 + col < n1_end) part[c0 + col] = s;
    }
}

// acht_n1_loc of the occupied bands n1 < nv for column j over G-vectors
// [split*chunk, +chunk): sum_ig aqsn[n1,ig] p2[j,ig] where the occupied solver
// took branch 1 (sch = P, else 0), bands in groups of 8 held in registers,
// block-wide fixed-order sums, weighted by 0.5 vcoul[igp(j)]
// conj(aqsm[n1,igp(j)]) and written to part[n1] (nv entries).
__device__ __forceinline__ void acht_occ_item(const cplx *__restrict__ p2, const uint8_t *__restrict__ br1,
                                              const cplx *__restrict__ aqsn, const cplx *__restrict__ bg,
                                              const int32_t *__restrict__ idx, const double *__restrict__ vcoul,
                                              int ngpown, int ncouls, int nv, int chunk, int j, int split,
                                              cplx *__restrict__ part, cplx *scratch)
{
    constexpr int G = 8;
    const int ig_begin = split * chunk;
    const int ig_end = ig_begin + chunk < ncouls ? ig_begin + chunk : ncouls;
    const size_t row = (size_t)j * ncouls;
    const double w = 0.5 * vcoul[igp_of(idx, j, ncouls)];
    for (int n0 = 0; n0 < nv; n0 += G) {
        cplx acc[G];
#pragma unroll
        for (int u = 0; u < G; ++u) acc[u] = czero();
        for (int ig = ig_begin + threadIdx.x; ig < ig_end; ig += blockDim.x) {
            if (!br1[row + ig]) continue;
            const cplx q = p2[row + ig];
#pragma unroll
            for (int u = 0; u < G; ++u)
                if (n0 + u < nv) acc[u] = cfma(aqsn[(size_t)(n0 + u) * ncouls + ig], q, acc[u]);
        }
#pragma unroll
        for (int u = 0; u < G; ++u) {
            if (n0 + u >= nv) break;  // uniform
            const cplx s = block_sum(acc[u], scratch);
            if (threadIdx.x == 0)
                part[n0 + u] = cscale(cmul_fast(cconj(bg[(size_t)(n0 + u) * ngpown + j]), s), w);
        }
    }
}

// Stage 4: block `part` < nfreq forms achtemp[part] / asxtemp[part] as
// fixed-order sums over the (column, split) partials; block part == nfreq
// sets acht_n1_loc to NaN (the factored forms write it in a later stage when
// it is requested) and sums the per-(band, column group) achstemp partials.
__device__ __forceinline__ void final_block(const cplx *col_ach, const cplx *col_asx, const cplx *part_achs, int ncols,
                                            int nfreq, int nb, int nv, cplx *out, cplx *scratch, int part)
{
    if (part < nfreq) {
        const int iw = part;
        cplx s = czero(), x = czero();
        for (int c = threadIdx.x; c < ncols; c += blockDim.x) {
            s = cadd(s, col_ach[(size_t)c * nfreq + iw]);
            x = cadd(x, col_asx[(size_t)c * nfreq + iw]);
        }
        s = block_sum(s, scratch);
        x = block_sum(x, scratch);
        if (threadIdx.x == 0) {
            out[iw] = s;
            out[nfreq + iw] = nv > 0 ? x : czero();
        }
        return;
    }
    const double nan = __builtin_nan("");
    for (int n1 = threadIdx.x; n1 < nb; n1 += blockDim.x) out[2 * nfreq + n1] = cplx{nan, nan};
    cplx h = czero();
    for (int t = threadIdx.x; t < nb * kAchSplit; t += blockDim.x) h = cadd(h, part_achs[t]);
    h = block_sum(h, scratch);
    if (threadIdx.x == 0) out[2 * nfreq + nb] = h;
}

}  // namespace fact
}  // namespace gpp
