// gpp_factored_dev.hpp -- device building blocks of the factored form, shared
// by the single-instance kernels (gpp_factored.hip, one grid per stage) and
// the batched persistent kernels (gpp_batch.hip, one work list per stage).
//
// Factorisation (wx_array is band-independent, gppComplex.cpp:323-329):
//   Mu[j,ig] = sum_{n1 >= nv} conj(aqsm[n1,igp(j)]) * aqsn[n1,ig]
//   Mo[j,ig] = sum_{n1 <  nv} conj(aqsm[n1,igp(j)]) * aqsn[n1,ig]
//   achtemp[iw] = sum_j 0.5 vcoul[igp(j)] sum_ig (P*Mu + sch*Mo)
//   asxtemp[iw] = sum_j occ vcoul[igp(j)] sum_ig ssx*Mo
// with P = delw*I_eps (noflagOCC_solver, :139-146) and sch, ssx from
// ssxt_scht_solver (:16-63); sch = P in branch 1 and 0 otherwise.
#pragma once

#include "gpp_internal.hpp"

namespace gpp {
namespace fact {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kKS = 8;             // bands per LDS stage of the FP64 Gram tile (two MFMA k-steps)
constexpr int kWJ = 64, kWI = 32;  // FP64 Gram wave tile: 64 G' columns x 32 G-vectors
constexpr int kRedThreads = 256;   // pole-reduce and final-reduce blocks

// FP64 Gram block tile made of WJN x WIN waves.
template <int WJN, int WIN> struct GramCfg {
    static constexpr int TJ = kWJ * WJN, TI = kWI * WIN, NT = 64 * WJN * WIN;
};

// Stage 1: Bg[n1][j] = aqsm[n1][igp(j)] for element t of [nb x ngpown].
__device__ __forceinline__ void gather_elem(const cplx *aqsm, const int32_t *idx, int ncouls, int ngpown, size_t t,
                                            cplx *bg)
{
    const int n1 = (int)(t / ngpown), j = (int)(t % ngpown);
    bg[t] = aqsm[(size_t)n1 * ncouls + igp_of(idx, j, ncouls)];
}

// Stage 2: one TJ x TI tile of Mu = conj(Bg)^T * aqsn over bands [k_begin, k_end)
// on v_mfma_f64_16x16x4f64.  Lane layout (measured on gfx950): A[l%16][l/16],
// B[l/16][l%16], result r of lane l = C[4r + l/16][l%16]; MFMA rows are G'
// columns j and MFMA columns are G-vectors ig, so each result row is
// contiguous in ig.  sB/sA are the block's double-buffered LDS stages.
template <int WJN, int WIN>
__device__ __forceinline__ void gram_f64_tile(const cplx *__restrict__ bg, const cplx *__restrict__ aqsn, int ngpown,
                                              int ncouls, int k_begin, int k_end, int j0, int i0,
                                              cplx (*sB)[kKS][GramCfg<WJN, WIN>::TJ],
                                              cplx (*sA)[kKS][GramCfg<WJN, WIN>::TI], cplx *__restrict__ mout)
{
    typedef GramCfg<WJN, WIN> C;
    constexpr int TJ = C::TJ, TI = C::TI, NT = C::NT;
    constexpr int PB = kKS * TJ / NT, PA = kKS * TI / NT;  // staged elements per thread
    static_assert(PB * NT == kKS * TJ && PA * NT == kKS * TI, "stage must split evenly");
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    const int wj = (wave / WIN) * kWJ;
    const int wi = (wave % WIN) * kWI;

    v4d acc_re[4][2], acc_im[4][2];
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            acc_re[s][t] = v4d{0, 0, 0, 0};
            acc_im[s][t] = v4d{0, 0, 0, 0};
        }

    cplx stB[PB], stA[PA];
    auto load_stage = [&](int k0) {
#pragma unroll
        for (int q = 0; q < PB; ++q) {
            const int e = threadIdx.x + q * NT;
            const int kr = e / TJ, c = e % TJ, k = k0 + kr, j = j0 + c;
            stB[q] = (k < k_end && j < ngpown) ? bg[(size_t)k * ngpown + j] : czero();
        }
#pragma unroll
        for (int q = 0; q < PA; ++q) {
            const int e = threadIdx.x + q * NT;
            const int kr = e / TI, c = e % TI, k = k0 + kr, ig = i0 + c;
            stA[q] = (k < k_end && ig < ncouls) ? aqsn[(size_t)k * ncouls + ig] : czero();
        }
    };
    auto store_stage = [&](int buf) {
#pragma unroll
        for (int q = 0; q < PB; ++q) {
            const int e = threadIdx.x + q * NT;
            sB[buf][e / TJ][e % TJ] = stB[q];
        }
#pragma unroll
        for (int q = 0; q < PA; ++q) {
            const int e = threadIdx.x + q * NT;
            sA[buf][e / TI][e % TI] = stA[q];
        }
    };

    const int nst = (k_end - k_begin + kKS - 1) / kKS;
    load_stage(k_begin);
    store_stage(0);
    __syncthreads();
    for (int st = 0; st < nst; ++st) {
        const int buf = st & 1;
        if (st + 1 < nst) load_stage(k_begin + (st + 1) * kKS);
#pragma unroll
        for (int kk = 0; kk < kKS; kk += 4) {
            const int kr = kk + lane / 16;
            double bre[4], bim[4], bneg[4], are[2], aim[2];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const cplx b = sB[buf][kr][wj + 16 * s + lane % 16];
                bre[s] = b.re;
                bim[s] = b.im;
                bneg[s] = -b.im;
            }
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const cplx a = sA[buf][kr][wi + 16 * t + lane % 16];
                are[t] = a.re;
                aim[t] = a.im;
            }
            // conj(b) * a: re = br*ar + bi*ai, im = br*ai - bi*ar
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    acc_re[s][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(bre[s], are[t], acc_re[s][t], 0, 0, 0);
                    acc_re[s][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(bim[s], aim[t], acc_re[s][t], 0, 0, 0);
                    acc_im[s][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(bre[s], aim[t], acc_im[s][t], 0, 0, 0);
                    acc_im[s][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(bneg[s], are[t], acc_im[s][t], 0, 0, 0);
                }
        }
        if (st + 1 < nst) store_stage(buf ^ 1);
        __syncthreads();
    }

#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int j = j0 + wj + 16 * s + 4 * r + lane / 16;
                const int ig = i0 + wi + 16 * t + lane % 16;
                if (j < ngpown && ig < ncouls) mout[(size_t)j * ncouls + ig] = cplx{acc_re[s][t][r], acc_im[s][t][r]};
            }
}

// One instance's inputs and workspace as seen by the reduce stages.
struct ReduceArgs {
    const cplx *mu;  // [ngpown][ncouls]
    const cplx *ieps, *wtilde, *aqsn;
    const cplx *bg;  // [nb][ngpown]
    const int32_t *idx;
    const double *vcoul, *wx;
    int nv, ncouls, ngpown, nfreq, nsplit, chunk;
    double occ;
    cplx *col_ach, *col_asx;  // [ngpown][nsplit][nfreq]
};

// Stage 3: pole terms of column j over G-vectors [split*chunk, +chunk), for
// frequencies [iw0, iw0+FC).  Per element and frequency: delw =
// wt*conj(wdiff)*rden and P = delw*ie in the reference's order
// (gppComplex.cpp:139-146, no contraction; rden is the refined reciprocal,
// equal to the IEEE quotient), so the occupied branch tests (:35, :43) see
// the reference's values; sch = P in branch 1 (:37) shares the unoccupied
// product, ssx follows :38-51.  Block-wide fixed-order sums; writes the
// column partials (0.5*sum)*vcoul and sum*(occ*vcoul).  scratch: 4 entries.
template <int FC>
__device__ __forceinline__ void pole_reduce_item(const ReduceArgs &a, int j, int split, int iw0, cplx *scratch)
{
#pragma clang fp contract(off)
    const int ncouls = a.ncouls;
    const int nv = a.nv;
    const int ig_begin = split * a.chunk;
    const int ig_end = ig_begin + a.chunk < ncouls ? ig_begin + a.chunk : ncouls;
    const double limitone = 1.0 / (kTo1 * 4.0);
    double wxv[FC], wx2[FC], rwx4[FC];
#pragma unroll
    for (int f = 0; f < FC; ++f) {
        wxv[f] = (iw0 + f < a.nfreq) ? a.wx[iw0 + f] : 1.0;
        wx2[f] = wxv[f] * wxv[f];
        rwx4[f] = 1.00 / (wx2[f] * wx2[f]);
    }
    cplx sach[FC], sasx[FC];
#pragma unroll
    for (int f = 0; f < FC; ++f) sach[f] = sasx[f] = czero();

    const size_t row = (size_t)j * ncouls;
    for (int ig = ig_begin + threadIdx.x; ig < ig_end; ig += blockDim.x) {
        const cplx wt = a.wtilde[row + ig];
        const cplx ie = a.ieps[row + ig];
        const cplx mu = a.mu[row + ig];
        cplx mo = czero(), om2 = czero();
        if (nv > 0) {
            for (int n1 = 0; n1 < nv; ++n1)  // occupied bands, in band order
                mo = cadd(mo, cmul_fast(cconj(a.bg[(size_t)n1 * a.ngpown + j]), a.aqsn[(size_t)n1 * ncouls + ig]));
            om2 = cmul(csquare(wt), ie);  // Omega2 = wtilde^2 * I_eps, flagOCC_solver :117-118
        }
#pragma unroll
        for (int f = 0; f < FC; ++f) {
            const cplx wdiff = cdminus(wxv[f], wt);
            const double wdiffr = cmul(wdiff, cconj(wdiff)).re;
            const double rden = rcp_refined(wdiffr);
            const cplx delw = cscale(cmul(wt, cconj(wdiff)), rden);
            const cplx p = cmul(delw, ie);
            cplx m = mu;
            if (nv > 0) {
                const double delwr = cmul(delw, cconj(delw)).re;
                cplx ssx = czero();
                if ((wdiffr > 0.25) && (delwr < limitone)) {
                    m = cadd(mu, mo);  // sch = P for the occupied bands
                    ssx = cscale2(om2, wx2[f], rwx4[f]);
                } else if (delwr > kTo1) {
                    const cplx cden = cscale(cmul(csquare(wt), cdplus(0.50, delw)), 4.00);
                    const double rd = rcp_refined(cmul(cden, cconj(cden)).re);
                    ssx = cmul(cmul(cplx{-om2.re, -om2.im}, cconj(cden)), cscale(delw, rd));
                }
                if (wxv[f] < 0.00 && cabs_(ssx) > cabs_(ie) * kSexcut) ssx = czero();
                const cplx x = cmul_fast(ssx, mo);
                sasx[f] = cadd(sasx[f], x);
            }
            const cplx t = cmul_fast(p, m);
            sach[f] = cadd(sach[f], t);
        }
    }
    const double vc = a.vcoul[igp_of(a.idx, j, ncouls)];
#pragma unroll
    for (int f = 0; f < FC; ++f) {
        const cplx s = block_sum(sach[f], scratch);
        const cplx x = block_sum(sasx[f], scratch);
        if (threadIdx.x == 0 && iw0 + f < a.nfreq) {
            // (0.5 * sum) * vcoul and sum * (occ * vcoul), gppComplex.cpp:361-363, :379
            const size_t o = ((size_t)j * a.nsplit + split) * a.nfreq + iw0 + f;
            a.col_ach[o] = cscale(cscale(s, 0.5), vc);
            a.col_asx[o] = cscale(x, a.occ * vc);
        }
    }
}

// Stage 4 (one block): achtemp/asxtemp as fixed-order sums over the
// (column, split) partials, acht_n1_loc set to NaN (not produced by this
// form), achstemp as the fixed-order sum of the per-band partials.
__device__ __forceinline__ void final_block(const cplx *col_ach, const cplx *col_asx, const cplx *part_achs, int ncols,
                                            int nfreq, int nb, int nv, cplx *out, cplx *scratch)
{
    for (int iw = 0; iw < nfreq; ++iw) {
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
    }
    const double nan = __builtin_nan("");
    for (int n1 = threadIdx.x; n1 < nb; n1 += blockDim.x) out[2 * nfreq + n1] = cplx{nan, nan};
    cplx h = czero();
    for (int n1 = threadIdx.x; n1 < nb; n1 += blockDim.x) h = cadd(h, part_achs[n1]);
    h = block_sum(h, scratch);
    if (threadIdx.x == 0) out[2 * nfreq + nb] = h;
}

}  // namespace fact
}  // namespace gpp
