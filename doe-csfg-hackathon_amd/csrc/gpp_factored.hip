// gpp_factored.hip -- factored form of the GPP sigma band loop on gfx950.
//
// The frequency grid wx_array is band-independent (gppComplex.cpp:323-329
// computes it once, before the band loop), so the pole factors
//   P(iw,j,ig)   = delw * I_eps              (noflagOCC_solver, :143-146)
//   sch, ssx     (iw,j,ig)                  (ssxt_scht_solver, :16-63)
// do not depend on the band n1, and the band sum factors through the
// complex Gram products
//   Mu[j,ig] = sum_{n1 >= nv} conj(aqsm[n1,igp(j)]) * aqsn[n1,ig]
//   Mo[j,ig] = sum_{n1 <  nv} conj(aqsm[n1,igp(j)]) * aqsn[n1,ig]
// giving
//   achtemp[iw] = sum_j 0.5 vcoul[igp(j)] sum_ig (P*Mu + sch*Mo)
//   asxtemp[iw] = sum_j occ vcoul[igp(j)] sum_ig ssx*Mo.
// (For n1 < nv the reference's sch equals P in branch 1 and 0 otherwise.)
//
// Stages:
//   1. k_gather_aqsm      Bg[n1][j] = aqsm[n1][igp(j)]          (13 MB at 512x1638)
//   2. k_gram_f64         Mu = conj(Bg)^T * aqsn on FP64 MFMA (v_mfma_f64_16x16x4),
//                         LDS-staged 128x128 tiles, written to HBM as [ngpown][ncouls]
//   3. k_pole_reduce      streams wtilde, I_eps, Mu (and the nv occupied aqsn rows),
//                         evaluates the pole factors in FP64 (reference order and IEEE
//                         division on the occupied branches) and reduces per column j
//   4. fixed-order column reduction + achstemp (reduce_achstemp, :66-101).
// acht_n1_loc needs a per-band sum and is not produced by this form.
#include "gpp_internal.hpp"

namespace gpp {
namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kTJ = 128;      // Gram tile: G' columns (MFMA rows)
constexpr int kTI = 128;      // Gram tile: G-vectors (MFMA columns)
constexpr int kKS = 8;        // bands per LDS stage (two MFMA k-steps of 4)
constexpr int kGramThreads = 512;  // 8 waves: 2 (j) x 4 (ig), wave tile 64 x 32
constexpr int kWJ = 64, kWI = 32;  // wave tile
constexpr int kRedThreads = 256;
constexpr int kRedChunk = 8192;    // G-vectors per pole-reduce block

__host__ inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct FactLayout {
    size_t bg, m, col_ach, col_asx, part_achs, total;
};

FactLayout fact_layout(const DevProblem &p)
{
    FactLayout L;
    size_t off = 0;
    L.bg = off;        off += align256(sizeof(cplx) * (size_t)p.nb * p.ngpown);
    L.m = off;         off += align256(sizeof(cplx) * (size_t)p.ngpown * p.ncouls);
    const size_t nsplit = (size_t)(p.ncouls + kRedChunk - 1) / kRedChunk;
    L.col_ach = off;   off += align256(sizeof(cplx) * (size_t)p.ngpown * nsplit * p.nfreq);
    L.col_asx = off;   off += align256(sizeof(cplx) * (size_t)p.ngpown * nsplit * p.nfreq);
    L.part_achs = off; off += align256(sizeof(cplx) * ((size_t)p.nb + 1));
    L.total = off;
    return L;
}

// Stage 1: Bg[n1][j] = aqsm[n1][igp(j)].
__global__ void k_gather_aqsm(const cplx *__restrict__ aqsm, const int32_t *__restrict__ idx, int nb, int ncouls,
                              int ngpown, cplx *__restrict__ bg)
{
    const size_t n = (size_t)nb * ngpown;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        const int n1 = (int)(t / ngpown), j = (int)(t % ngpown);
        bg[t] = aqsm[(size_t)n1 * ncouls + igp_of(idx, j, ncouls)];
    }
}

// Stage 2: Mu[j][ig] = sum_{k in [k_begin, k_end)} conj(Bg[k][j]) * aqsn[k][ig].
// Grid (ceil(ngpown/128), ceil(ncouls/128)); block 512 threads.
// MFMA f64 16x16x4 operand layout (lane l): A[i=l%16][k=l/16], B[k=l/16][j=l%16],
// result element r of lane l = C[4r + l/16][l%16].  Here MFMA rows are G'
// columns j, MFMA columns are G-vectors ig, so each result row is contiguous in ig.
__global__ __launch_bounds__(kGramThreads) void k_gram_f64(const cplx *__restrict__ bg, const cplx *__restrict__ aqsn,
                                                           int ngpown, int ncouls, int k_begin, int k_end,
                                                           cplx *__restrict__ mout)
{
    __shared__ cplx sB[2][kKS][kTJ];  // conj operand rows (bands) x columns j
    __shared__ cplx sA[2][kKS][kTI];  // aqsn rows (bands) x G-vectors
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    const int wj = (wave / 4) * kWJ;  // wave's j offset in the tile (0 or 64)
    const int wi = (wave % 4) * kWI;  // wave's ig offset in the tile (0..96)
    const int j0 = blockIdx.x * kTJ;  // j tiles vary fastest: consecutive blocks share one aqsn tile
    const int i0 = blockIdx.y * kTI;

    v4d acc_re[4][2], acc_im[4][2];
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            acc_re[s][t] = v4d{0, 0, 0, 0};
            acc_im[s][t] = v4d{0, 0, 0, 0};
        }

    // Stage loader: 2 * 8 * 128 complex per stage, 2 per thread per array.
    cplx stB[2], stA[2];
    auto load_stage = [&](int k0) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int e = threadIdx.x + q * kGramThreads;  // 0..1023
            const int kr = e / kTJ, c = e % kTJ;
            const int k = k0 + kr;
            const int j = j0 + c, ig = i0 + c;
            stB[q] = (k < k_end && j < ngpown) ? bg[(size_t)k * ngpown + j] : czero();
            stA[q] = (k < k_end && ig < ncouls) ? aqsn[(size_t)k * ncouls + ig] : czero();
        }
    };
    auto store_stage = [&](int buf) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int e = threadIdx.x + q * kGramThreads;
            const int kr = e / kTJ, c = e % kTJ;
            sB[buf][kr][c] = stB[q];
            sA[buf][kr][c] = stA[q];
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

struct ReduceArgs {
    const cplx *mu;      // [ngpown][ncouls]
    const cplx *ieps, *wtilde, *aqsn;
    const cplx *bg;      // [nb][ngpown]
    const int32_t *idx;
    const double *vcoul, *wx;
    int nv, ncouls, ngpown, nfreq, nsplit;
    double occ;
    cplx *col_ach, *col_asx;  // [ngpown][nsplit][nfreq]
};

// Stage 3: block (j, split, frequency chunk) reduces the pole terms of column j
// over G-vectors [split*kRedChunk, +kRedChunk).  Per element and frequency:
//   delw = wt*conj(wdiff)*rden and P = delw*ie in the reference's order
//   (gppComplex.cpp:139-146, no contraction; rden = the refined reciprocal, equal
//   to the IEEE quotient), so the occupied branch tests (:35, :43) see the
//   reference's values; sch = P in branch 1 (:37) is shared with the
//   unoccupied term, and ssx follows :38-51.
template <int FC>
__global__ __launch_bounds__(kRedThreads) void k_pole_reduce(ReduceArgs a)
{
#pragma clang fp contract(off)
    __shared__ cplx scratch[kRedThreads / kWave];
    const int j = blockIdx.x;
    const int split = blockIdx.y;
    const int iw0 = blockIdx.z * FC;
    const int ncouls = a.ncouls;
    const int nv = a.nv;
    const int ig_begin = split * kRedChunk;
    const int ig_end = ig_begin + kRedChunk < ncouls ? ig_begin + kRedChunk : ncouls;
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
                    m = cadd(mu, mo);                  // sch = P for the occupied bands
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

// Stage 4: achtemp/asxtemp over columns (fixed order), achstemp over bands.
__global__ __launch_bounds__(kRedThreads) void k_fact_final(const cplx *col_ach, const cplx *col_asx,
                                                            const cplx *part_achs, int ncols, int nfreq, int nb,
                                                            int nv, cplx *out)
{
    __shared__ cplx scratch[kRedThreads / kWave];
    for (int iw = 0; iw < nfreq; ++iw) {
        cplx s = czero(), x = czero();
        for (int c = threadIdx.x; c < ncols; c += blockDim.x) {  // (column, split) pairs in order
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

template <int FC>
void launch_reduce_fc(const ReduceArgs &a, hipStream_t s)
{
    dim3 grid(a.ngpown, a.nsplit, (a.nfreq + FC - 1) / FC);
    hipLaunchKernelGGL((k_pole_reduce<FC>), grid, dim3(kRedThreads), 0, s, a);
}

#define GPP_TRY(x)                       \
    do {                                 \
        hipError_t e_ = (x);             \
        if (e_ != hipSuccess) return e_; \
    } while (0)

}  // namespace

size_t factored_workspace_bytes(const DevProblem &p, GramKind kind)
{
    size_t n = fact_layout(p).total;
    if (kind.ozaki_slices) n += ozaki_workspace_bytes(p, kind.ozaki_slices, kind.int8);
    return n;
}

hipError_t run_factored(const DevProblem &p, GramKind kind, void *work, const DevResults &res, hipStream_t s,
                        StageTimer *timer)
{
    const FactLayout L = fact_layout(p);
    char *w = static_cast<char *>(work);
    cplx *bg = reinterpret_cast<cplx *>(w + L.bg);
    cplx *mu = reinterpret_cast<cplx *>(w + L.m);
    cplx *col_ach = reinterpret_cast<cplx *>(w + L.col_ach);
    cplx *col_asx = reinterpret_cast<cplx *>(w + L.col_asx);
    cplx *part_achs = reinterpret_cast<cplx *>(w + L.part_achs);

    if (timer) timer->begin(s);
    {
        const size_t n = (size_t)p.nb * p.ngpown;
        size_t blocks = (n + 255) / 256;
        if (blocks > 4096) blocks = 4096;
        hipLaunchKernelGGL(k_gather_aqsm, dim3((unsigned)blocks), dim3(256), 0, s, p.aqsm, p.idx, p.nb, p.ncouls,
                           p.ngpown, bg);
        GPP_TRY(hipGetLastError());
    }
    GPP_TRY(launch_achstemp(p, part_achs, s));
    if (timer) timer->mark(s, "factored_gather_achstemp");
    if (kind.ozaki_slices) {
        GPP_TRY(launch_gram_ozaki(p, kind.ozaki_slices, kind.int8, kind.layout, bg, w + L.total, mu, s));
        if (timer) timer->mark(s, kind.int8 ? "factored_gram_ozaki_i8" : "factored_gram_ozaki");
    } else {
        dim3 grid((p.ngpown + kTJ - 1) / kTJ, (p.ncouls + kTI - 1) / kTI);
        hipLaunchKernelGGL(k_gram_f64, grid, dim3(kGramThreads), 0, s, bg, p.aqsn, p.ngpown, p.ncouls, p.nv, p.nb, mu);
        GPP_TRY(hipGetLastError());
        if (timer) timer->mark(s, "factored_gram_f64");
    }
    ReduceArgs ra;
    ra.mu = mu; ra.ieps = p.ieps; ra.wtilde = p.wtilde; ra.aqsn = p.aqsn; ra.bg = bg; ra.idx = p.idx;
    ra.vcoul = p.vcoul; ra.wx = p.wx; ra.nv = p.nv; ra.ncouls = p.ncouls; ra.ngpown = p.ngpown;
    ra.nfreq = p.nfreq; ra.occ = p.occ; ra.col_ach = col_ach; ra.col_asx = col_asx;
    ra.nsplit = (p.ncouls + kRedChunk - 1) / kRedChunk;
    switch (p.nfreq) {
    case 1: launch_reduce_fc<1>(ra, s); break;
    case 2: launch_reduce_fc<2>(ra, s); break;
    case 3: launch_reduce_fc<3>(ra, s); break;
    case 4: launch_reduce_fc<4>(ra, s); break;
    default: launch_reduce_fc<8>(ra, s); break;
    }
    GPP_TRY(hipGetLastError());
    if (timer) timer->mark(s, "factored_pole_reduce");
    hipLaunchKernelGGL(k_fact_final, dim3(1), dim3(kRedThreads), 0, s, col_ach, col_asx, part_achs,
                       p.ngpown * ra.nsplit, p.nfreq, p.nb, p.nv, res.base);
    GPP_TRY(hipGetLastError());
    if (timer) timer->mark(s, "factored_final");
    return hipSuccess;
}

}  // namespace gpp
