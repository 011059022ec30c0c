// gpp_factored.hip -- factored form of the GPP sigma band loop on gfx950,
// one problem per call (the batched persistent form is gpp_batch.hip).
//
// Stages (device bodies in gpp_factored_dev.hpp):
//   1. k_achstemp      Bg[n1][j] = aqsm[n1][igp(j)] (13 MB at 512x1638) and the
//                      achstemp band sums (reduce_achstemp, gppComplex.cpp:66-101)
//   2. Gram product    Mu = conj(Bg)^T * aqsn over bands [nv, nb):
//        FP64 MFMA (k_gram_f64, tile chosen by GramKind::tile), or the Ozaki
//        split of gpp_ozaki.hip; written to HBM as [ngpown][ncouls]
//   3. k_pole_reduce   streams wtilde, I_eps, Mu (and the nv occupied aqsn rows),
//                      evaluates the pole factors in FP64, reduces per column
//   4. k_fact_final    fixed-order column reduction; achstemp (reduce_achstemp,
//                      gppComplex.cpp:66-101) from k_achstemp.
// When acht_n1_loc is requested (GramKind::acht), stage 3 also writes the
// pole factor p2 of frequency 2 and, with two occupied bands (and no negative
// frequency), sums the occupied bands' terms (the masked p2 against their
// aqsn rows, already loaded for Mo); and
//   5. k_acht_f64      second FP64 MFMA product Y = P2 * aqsn^T for the
//                      unoccupied bands, split over G-vectors, reduced per
//                      band in its epilogue; for the other occupied band
//                      counts k_acht_occ streams the masked P2 (stage 3
//                      writes the branch mask) against the occupied aqsn rows
//   6. k_acht_final    fixed-order sum of the per-band partials.
#include <string.h>

#include <algorithm>

#include "gpp_factored_dev.hpp"

namespace gpp {
namespace {

using namespace fact;

__host__ inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// acht product tiles: rows j in 128s for both band ranges; unoccupied bands
// in columns of 128, occupied bands (few) in columns of 32.
typedef AchtCfg<2, 4> AchtU;

struct AchtSplit {
    int jtiles, nks, chunk;
};

// Split of the G-vector sum so that the unoccupied launch has ~4096 blocks
// (C1: 3.9 % faster than ~2048, 7 % faster than ~1024).
AchtSplit acht_split(const DevProblem &p)
{
    AchtSplit a;
    a.jtiles = (p.ngpown + AchtU::TJ - 1) / AchtU::TJ;
    const int ntiles = (p.nb - p.nv + AchtU::TN - 1) / AchtU::TN;
    const int tiles = a.jtiles * (ntiles > 0 ? ntiles : 1);
    int nks = (4096 + tiles - 1) / tiles;
    const int max_nks = p.ncouls / 512 > 1 ? p.ncouls / 512 : 1;
    if (nks > max_nks) nks = max_nks;
    a.chunk = ((p.ncouls + nks - 1) / nks + kAKS - 1) / kAKS * kAKS;
    a.nks = (p.ncouls + a.chunk - 1) / a.chunk;
    return a;
}

// The fused Gram + pole reduce (k_gram_pole) serves at most three
// frequencies, none negative, at most kFuseNv occupied bands, without
// acht_n1_loc, on the 128 x 128 FP64 tile.
typedef GramCfg<2, 4> FuseCfg;

bool fused_reduce(const DevProblem &p, const GramKind &k)
{
    return k.fuse && k.ozaki_slices == 0 && k.tile == 0 && !k.acht && p.nfreq <= 3 && !p.wx_neg &&
           p.nv <= kFuseNv;
}

struct FactLayout {
    size_t bg, m, col_ach, col_asx, col2_ach, col2_asx, part_achs, p2, br1, acht_part, acht_occ, total;
    int nsplit;  // G-vector ranges of the pole-reduce partials (the fused form: G-vector tiles)
    bool fused;
    AchtSplit as;
};

FactLayout fact_layout(const DevProblem &p, const GramKind &k)
{
    FactLayout L;
    L.fused = fused_reduce(p, k);
    L.nsplit = L.fused ? (p.ncouls + FuseCfg::TI - 1) / FuseCfg::TI : (p.ncouls + k.red_chunk - 1) / k.red_chunk;
    size_t off = 0;
    L.bg = off;        off += align256(sizeof(cplx) * (size_t)p.nb * p.ngpown);
    L.m = off;         off += L.fused ? 0 : align256(sizeof(cplx) * (size_t)p.ngpown * p.ncouls);
    L.col_ach = off;   off += align256(sizeof(cplx) * (size_t)p.ngpown * L.nsplit * p.nfreq);
    L.col_asx = off;   off += align256(sizeof(cplx) * (size_t)p.ngpown * L.nsplit * p.nfreq);
    L.col2_ach = L.col2_asx = off;
    if (L.fused) {
        L.col2_ach = off; off += align256(sizeof(cplx) * (size_t)p.ngpown * p.nfreq);
        L.col2_asx = off; off += align256(sizeof(cplx) * (size_t)p.ngpown * p.nfreq);
    }
    L.part_achs = off; off += align256(sizeof(cplx) * ach_region_entries(p.nb, p.ngpown));
    L.p2 = L.br1 = L.acht_part = L.acht_occ = off;
    L.as = AchtSplit{0, 0, 0};
    if (k.acht) {
        L.as = acht_split(p);
        L.p2 = off;        off += align256(sizeof(cplx) * (size_t)p.ngpown * p.ncouls);
        // branch-1 mask for k_acht_occ, where the reduce does not sum the
        // occupied bands itself
        L.br1 = off;       off += p.nv > 0 && !reduce_sums_acht_occ(p.nv, p.wx_neg) ? align256((size_t)p.ngpown * p.ncouls) : 0;
        L.acht_part = off; off += align256(sizeof(cplx) * (size_t)L.as.jtiles * L.as.nks * p.nb);
        L.acht_occ = off;  off += align256(sizeof(cplx) * (size_t)p.ngpown * L.nsplit * (p.nv > 0 ? p.nv : 1));
    }
    L.total = off;
    return L;
}

// 1-D grid over the ceil(ngpown/TJ) x ceil(ncouls/TI) tiles, G' tiles
// fastest in tile order and tiles spread over the XCDs by xcd_tile, so the
// blocks sharing one aqsn tile run on one XCD.  Every tile runs two waves per
// SIMD (at most 256 registers per lane).
template <int WJN, int WIN>
__global__ __launch_bounds__(64 * WJN * WIN, WJN * WIN >= 4 ? 8 / (WJN * WIN) : 2) void k_gram_f64(const cplx *__restrict__ bg,
                                                                    const cplx *__restrict__ aqsn, int ngpown,
                                                                    int ncouls, int k_begin, int k_end,
                                                                    cplx *__restrict__ mout)
{
    typedef GramCfg<WJN, WIN> C;
    __shared__ cplx sB[2][C::KS][C::TJ];
    __shared__ cplx sA[2][C::KS][C::TI];
    const int nj = (ngpown + C::TJ - 1) / C::TJ;
    const int ni = (ncouls + C::TI - 1) / C::TI;
    const int t = xcd_tile(blockIdx.x, nj * ni);
    gram_f64_tile<WJN, WIN>(bg, aqsn, ngpown, ncouls, k_begin, k_end, (t % nj) * C::TJ, (t / nj) * C::TI, sB, sA,
                            mout);
}

// Stages 2 and 3 fused (at most three frequencies, none negative,
// acht_n1_loc not requested): the Gram tile of k_gram_f64 in the transposed
// orientation, then the pole terms of its elements summed per column in the
// epilogue (gram_pole_epilogue), so Mu never reaches HBM.  Same grid and tile
// order as k_gram_f64; col_ach/col_asx get one sum per (column, G-vector
// tile, frequency).
template <int F>
__global__ __launch_bounds__(512, 1) void k_gram_pole(PoleTileArgs a)
{
    typedef GramCfg<2, 4> C;
    __shared__ cplx lds[FuseLds<F>::kLds];
    auto *sB = reinterpret_cast<cplx(*)[C::KS][C::TJ]>(lds);
    auto *sA = reinterpret_cast<cplx(*)[C::KS][C::TI]>(lds + 2 * C::KS * C::TJ);
    const int nj = (a.ngpown + C::TJ - 1) / C::TJ;
    const int ni = (a.ncouls + C::TI - 1) / C::TI;
    const int t = xcd_tile(blockIdx.x, nj * ni);
    const int j0 = (t % nj) * C::TJ, itile = t / nj, i0 = itile * C::TI;
    v4d acc1[4][2], acc2[4][2], acc3[4][2];
    gram_f64_acc<2, 4, true>(a.bg, a.aqsn, a.ngpown, a.ncouls, a.nv, a.nb, j0, i0, sB, sA, acc1, acc2, acc3);
    gram_pole_epilogue<F>(a, j0, i0, itile, acc1, acc2, acc3, lds);
}

// One wave per column: the fused epilogue's tile sums, summed and scaled.
__global__ __launch_bounds__(256) void k_pole_cols(const cplx *col_ach, const cplx *col_asx, int nsplit, int nfreq,
                                                   const int32_t *idx, const double *vcoul, int ncouls, int ngpown,
                                                   double occ, cplx *out_ach, cplx *out_asx)
{
    const int j = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (j < ngpown) pole_cols_wave(col_ach, col_asx, nsplit, nfreq, idx, vcoul, ncouls, occ, j, out_ach, out_asx);
}

// Grid (ngpown, nsplit, ceil(nfreq/FC)).  NV: occupied band count fixed at
// compile time (0: any).  Three waves per SIMD: with the wave-uniform pole
// inputs held in scalar registers the FC = 8 variants fit 168 VGPRs (the
// 8-occupied-band one spills 10), and at 16 frequencies x 8 occupied bands
// three waves run the reduce 12% faster than two.  Two for FC = 16 (its 64
// accumulators; the plain variants fit 208-241 VGPRs).  Four for the plain
// kernels of at most three frequencies: their waves wait on memory half of
// the time, and a fourth hides more of it despite 17 spilled registers
// (3 frequencies x 2 occupied bands: 0.510 -> 0.486 ms at 512/2/32768/1638);
// the p2-storing and negative-frequency variants spill 30-72 and keep three
// (the p2 one at four: 0.70 -> 1.28 ms at 512/2/32768/1638).
// The negative-frequency kernels of 8 frequencies per pass run two (their
// cutoff code spilled 84-186 registers at three).
template <int FC, int NV, bool P2, bool NEG = false>
__global__ __launch_bounds__(kRedThreads, (FC <= 3 && !P2 && !NEG) ? 4 : FC > 8 || (NEG && FC > 4) ? 2 : 3) void k_pole_reduce(ReduceArgs a)
{
    __shared__ cplx scratch[kRedThreads / kWave];
    const int iw_begin = blockIdx.z * a.fpb;
    const int iw_end = iw_begin + a.fpb < a.nfreq ? iw_begin + a.fpb : a.nfreq;
    pole_reduce_item<FC, NV, P2, (FC > 4 ? 4 : FC), NEG>(a, blockIdx.x, blockIdx.y, iw_begin, iw_end, scratch);
}

__global__ __launch_bounds__(kRedThreads) void k_fact_final(const cplx *col_ach, const cplx *col_asx,
                                                            const cplx *part_achs, int ncols, int nfreq, int nb,
                                                            int nv, cplx *out)
{
    __shared__ cplx scratch[kRedThreads / kWave];
    final_block(col_ach, col_asx, part_achs, ncols, nfreq, nb, nv, out, scratch, blockIdx.x);
}

// 1-D grid over (column tiles, j tiles, G-vector splits), column tiles
// fastest in tile order and spread over the XCDs by xcd_tile, so the blocks
// that share one R tile run on one XCD.  part: [jtiles * nks][nb].
template <int WJN, int WIN>
__global__ __launch_bounds__(64 * WJN * WIN) void k_acht_f64(const cplx *__restrict__ p2,
                                                                    const cplx *__restrict__ aqsn,
                                                                    const cplx *__restrict__ bg,
                                                                    const int32_t *__restrict__ idx,
                                                                    const double *__restrict__ vcoul, int ngpown,
                                                                    int ncouls, int nb, int n1_begin, int n1_end,
                                                                    int chunk, int nks, int jtiles,
                                                                    cplx *__restrict__ part)
{
    typedef AchtCfg<WJN, WIN> C;
    __shared__ cplx sR[2][2][C::TJ][kAK];
    __shared__ cplx sA[2][2][C::TN][kAK];
    __shared__ cplx red[WJN][C::TN];
    const int nn = (n1_end - n1_begin + C::TN - 1) / C::TN;
    const int t = xcd_tile(blockIdx.x, nn * jtiles * nks);
    const int ct = t % nn, jt = (t / nn) % jtiles, ks = t / (nn * jtiles);
    const int k_begin = ks * chunk;
    const int k_end = k_begin + chunk < ncouls ? k_begin + chunk : ncouls;
    acht_f64_tile<WJN, WIN>(p2, aqsn, bg, idx, vcoul, ngpown, ncouls, n1_end, k_begin, k_end,
                                  jt * C::TJ, n1_begin + ct * C::TN, sR, sA, red,
                                  part + ((size_t)jt * nks + ks) * nb);
}

// Grid (ngpown, nsplit): the occupied bands' acht_n1_loc partials (where the
// reduce does not sum them: reduce_sums_acht_occ).
__global__ __launch_bounds__(kRedThreads) void k_acht_occ(const cplx *__restrict__ p2,
                                                          const uint8_t *__restrict__ br1,
                                                          const cplx *__restrict__ aqsn, const cplx *__restrict__ bg,
                                                          const int32_t *__restrict__ idx,
                                                          const double *__restrict__ vcoul, int ngpown, int ncouls,
                                                          int nv, int chunk, int nsplit, cplx *__restrict__ part_occ)
{
    __shared__ cplx scratch[kRedThreads / kWave];
    acht_occ_item(p2, br1, aqsn, bg, idx, vcoul, ngpown, ncouls, nv, chunk, blockIdx.x, blockIdx.y,
                  part_occ + ((size_t)blockIdx.x * nsplit + blockIdx.y) * nv, scratch);
}

// acht_n1_loc[n1] as the fixed-order sum of its partials.  Blocks
// [0, ceil(nb/16)): 16 bands x 16 contiguous groups of the (j tile, split)
// partials of the product (bands n1 >= nv), groups combined in order.
// Blocks after them: one per occupied band n1 < nv, its (column, split)
// partials of k_acht_occ strided over the 256 threads, then block_sum.
constexpr int kAchtFinalBands = 16;
__global__ __launch_bounds__(256) void k_acht_final(const cplx *__restrict__ part, int nparts,
                                                    const cplx *__restrict__ part_occ, int nparts_occ, int nb,
                                                    int nv, cplx *__restrict__ out)
{
    __shared__ cplx grp[256 / kAchtFinalBands][kAchtFinalBands];
    const int nblk = (nb + kAchtFinalBands - 1) / kAchtFinalBands;
    if ((int)blockIdx.x < nblk) {
        const int n1 = blockIdx.x * kAchtFinalBands + threadIdx.x % kAchtFinalBands;
        const int g = threadIdx.x / kAchtFinalBands;
        constexpr int NG = 256 / kAchtFinalBands;
        cplx s = czero();
        if (n1 >= nv && n1 < nb) {
            const int q0 = (int)((long long)nparts * g / NG), q1 = (int)((long long)nparts * (g + 1) / NG);
            for (int q = q0; q < q1; ++q) s = cadd(s, part[(size_t)q * nb + n1]);
        }
        grp[g][threadIdx.x % kAchtFinalBands] = s;
        __syncthreads();
        if (g == 0 && n1 >= nv && n1 < nb) {
            cplx t = grp[0][threadIdx.x];
            for (int k = 1; k < NG; ++k) t = cadd(t, grp[k][threadIdx.x]);
            out[n1] = t;
        }
    } else {
        const int n1 = (int)blockIdx.x - nblk;  // < nv
        cplx s = czero();
        for (int q = threadIdx.x; q < nparts_occ; q += blockDim.x) s = cadd(s, part_occ[(size_t)q * nv + n1]);
        s = block_sum(s, &grp[0][0]);
        if (threadIdx.x == 0) out[n1] = s;
    }
}

template <int WJN, int WIN>
void launch_gram_cfg(const DevProblem &p, const cplx *bg, cplx *mu, hipStream_t s)
{
    typedef GramCfg<WJN, WIN> C;
    const unsigned tiles = (unsigned)((p.ngpown + C::TJ - 1) / C::TJ) * (unsigned)((p.ncouls + C::TI - 1) / C::TI);
    hipLaunchKernelGGL((k_gram_f64<WJN, WIN>), dim3(tiles), dim3(C::NT), 0, s, bg, p.aqsn, p.ngpown, p.ncouls, p.nv,
                       p.nb, mu);
}

// P2: also store the pole factor of frequency acht_iw and either sum the
// occupied bands' acht_n1_loc terms (reduce_sums_acht_occ) or store the
// occupied branch mask.  Problems with a negative frequency run the generic
// NEG kernel.
template <int FC, bool P2>
void launch_reduce_p2(const ReduceArgs &a, dim3 grid, hipStream_t s)
{
    if (a.neg) {  // at most 8 per pass (the 16-wide variant spills)
        hipLaunchKernelGGL((k_pole_reduce<(FC > 8 ? 8 : FC), 0, P2, true>), grid, dim3(kRedThreads), 0, s, a);
        return;
    }
    switch (a.nv) {  // the occupied band counts of the common shapes
    case 2: hipLaunchKernelGGL((k_pole_reduce<FC, 2, P2>), grid, dim3(kRedThreads), 0, s, a); break;
    case 4: hipLaunchKernelGGL((k_pole_reduce<FC, 4, P2>), grid, dim3(kRedThreads), 0, s, a); break;
    case 8: hipLaunchKernelGGL((k_pole_reduce<FC, 8, P2>), grid, dim3(kRedThreads), 0, s, a); break;
    default: hipLaunchKernelGGL((k_pole_reduce<FC, 0, P2>), grid, dim3(kRedThreads), 0, s, a); break;
    }
}

template <int FC>
void launch_reduce_fc(ReduceArgs a, hipStream_t s)
{
    if (a.fpb < FC) a.fpb = FC;
    dim3 grid(a.ngpown, a.nsplit, (a.nfreq + a.fpb - 1) / a.fpb);
    if (a.p2)
        launch_reduce_p2<FC, true>(a, grid, s);
    else
        launch_reduce_p2<FC, false>(a, grid, s);
}

template <int F>
void launch_gram_pole_f(const DevProblem &p, const PoleTileArgs &a, hipStream_t s)
{
    const unsigned tiles =
        (unsigned)((p.ngpown + FuseCfg::TJ - 1) / FuseCfg::TJ) * (unsigned)((p.ncouls + FuseCfg::TI - 1) / FuseCfg::TI);
    hipLaunchKernelGGL((k_gram_pole<F>), dim3(tiles), dim3(FuseCfg::NT), 0, s, a);
}

#define GPP_TRY(x)                       \
    do {                                 \
        hipError_t e_ = (x);             \
        if (e_ != hipSuccess) return e_; \
    } while (0)

}  // namespace

size_t factored_workspace_bytes(const DevProblem &p, GramKind kind)
{
    size_t n = fact_layout(p, kind).total;
    if (kind.ozaki_slices) n += ozaki_workspace_bytes(p, kind.ozaki_slices, kind.int8);
    return n;
}

hipError_t launch_gram_f64(const DevProblem &p, int tile, const cplx *bg, cplx *mu, hipStream_t s)
{
    switch (tile) {
    case 0: launch_gram_cfg<2, 4>(p, bg, mu, s); break;  // 128 x 128, 8 waves
    case 1: launch_gram_cfg<1, 4>(p, bg, mu, s); break;  // 64 x 128, 4 waves
    case 2: launch_gram_cfg<2, 2>(p, bg, mu, s); break;  // 128 x 64, 4 waves
    case 3: launch_gram_cfg<1, 2>(p, bg, mu, s); break;  // 64 x 64, 2 waves
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// red_fc 0: passes of 16 from 16 frequencies up, none negative (C3, 16
// frequencies: one pass, 7.19 -> 6.74 ms against two of 8), else 8.  With a
// negative frequency at most 8 (that kernel spills at 16).
hipError_t launch_pole_reduce(const ReduceArgs &ra, int red_fc, hipStream_t s)
{
    int fc = red_fc == 4 || red_fc == 8 || red_fc == 16 ? red_fc : ra.nfreq >= 16 && !ra.neg ? 16 : 8;
    if (ra.neg && fc > 8) fc = 8;
    switch (ra.nfreq < 4 ? ra.nfreq : fc) {
    case 1: launch_reduce_fc<1>(ra, s); break;
    case 2: launch_reduce_fc<2>(ra, s); break;
    case 3: launch_reduce_fc<3>(ra, s); break;
    case 4: launch_reduce_fc<4>(ra, s); break;
    case 16: launch_reduce_fc<16>(ra, s); break;
    default: launch_reduce_fc<8>(ra, s); break;
    }
    return hipGetLastError();
}

hipError_t run_factored(const DevProblem &p, GramKind kind, void *work, const DevResults &res, hipStream_t s,
                        StageTimer *timer)
{
    if (p.wx_stride != 0) return hipErrorInvalidValue;  // the band sum factors only for a shared wx row
    const FactLayout L = fact_layout(p, kind);
    char *w = static_cast<char *>(work);
    cplx *bg = reinterpret_cast<cplx *>(w + L.bg);
    cplx *mu = reinterpret_cast<cplx *>(w + L.m);
    cplx *col_ach = reinterpret_cast<cplx *>(w + L.col_ach);
    cplx *col_asx = reinterpret_cast<cplx *>(w + L.col_asx);
    cplx *part_achs = reinterpret_cast<cplx *>(w + L.part_achs);

    if (timer) timer->begin(s);
    GPP_TRY(launch_achstemp(p, part_achs, bg, s));  // also gathers Bg
    if (timer) timer->mark(s, "factored_gather_achstemp");
    if (L.fused) {
        PoleTileArgs pa;
        pa.aqsn = p.aqsn; pa.ieps = p.ieps; pa.wtilde = p.wtilde; pa.bg = bg; pa.wx = p.wx;
        pa.nb = p.nb; pa.nv = p.nv; pa.ncouls = p.ncouls; pa.ngpown = p.ngpown; pa.nsplit = L.nsplit;
        pa.col_ach = col_ach; pa.col_asx = col_asx;
        switch (p.nfreq) {
        case 1: launch_gram_pole_f<1>(p, pa, s); break;
        case 2: launch_gram_pole_f<2>(p, pa, s); break;
        default: launch_gram_pole_f<3>(p, pa, s); break;
        }
        GPP_TRY(hipGetLastError());
        if (timer) timer->mark(s, "factored_gram_pole");
        cplx *c2a = reinterpret_cast<cplx *>(w + L.col2_ach);
        cplx *c2x = reinterpret_cast<cplx *>(w + L.col2_asx);
        hipLaunchKernelGGL(k_pole_cols, dim3((p.ngpown + 3) / 4), dim3(256), 0, s, col_ach, col_asx, L.nsplit,
                           p.nfreq, p.idx, p.vcoul, p.ncouls, p.ngpown, p.occ, c2a, c2x);
        GPP_TRY(hipGetLastError());
        if (timer) timer->mark(s, "factored_pole_cols");
        hipLaunchKernelGGL(k_fact_final, dim3(p.nfreq + 1), dim3(kRedThreads), 0, s, c2a, c2x, part_achs, p.ngpown,
                           p.nfreq, p.nb, p.nv, res.base);
        GPP_TRY(hipGetLastError());
        if (timer) timer->mark(s, "factored_final");
        return hipSuccess;
    }
    if (kind.ozaki_slices) {
        GPP_TRY(launch_gram_ozaki(p, kind.ozaki_slices, kind.int8, kind.layout, kind.group, bg, w + L.total, mu, s));
        if (timer) timer->mark(s, kind.int8 ? "factored_gram_ozaki_i8" : "factored_gram_ozaki");
    } else {
        GPP_TRY(launch_gram_f64(p, kind.tile, bg, mu, s));
        if (timer) timer->mark(s, "factored_gram_f64");
    }
    ReduceArgs ra;
    ra.mu = mu; ra.ieps = p.ieps; ra.wtilde = p.wtilde; ra.aqsn = p.aqsn; ra.bg = bg; ra.idx = p.idx;
    ra.vcoul = p.vcoul; ra.wx = p.wx; ra.nv = p.nv; ra.ncouls = p.ncouls; ra.ngpown = p.ngpown;
    ra.nfreq = p.nfreq; ra.occ = p.occ; ra.col_ach = col_ach; ra.col_asx = col_asx;
    ra.nsplit = L.nsplit; ra.chunk = kind.red_chunk;
    ra.fpb = kind.red_fpb > 0 ? kind.red_fpb : p.nfreq;
    ra.p2 = kind.acht ? reinterpret_cast<cplx *>(w + L.p2) : nullptr;
    const bool occ_in_reduce = kind.acht && p.nv > 0 && reduce_sums_acht_occ(p.nv, p.wx_neg);
    ra.br1 = kind.acht && p.nv > 0 && !occ_in_reduce ? reinterpret_cast<uint8_t *>(w + L.br1) : nullptr;
    ra.part_occ = occ_in_reduce ? reinterpret_cast<cplx *>(w + L.acht_occ) : nullptr;
    ra.acht_iw = p.nfreq > 2 ? 2 : p.nfreq - 1;  // sch_array[2], gppComplex.cpp:381
    ra.neg = p.wx_neg;
    GPP_TRY(launch_pole_reduce(ra, kind.red_fc, s));
    if (timer) timer->mark(s, "factored_pole_reduce");
    hipLaunchKernelGGL(k_fact_final, dim3(p.nfreq + 1), dim3(kRedThreads), 0, s, col_ach, col_asx, part_achs,
                       p.ngpown * L.nsplit, p.nfreq, p.nb, p.nv, res.base);
    GPP_TRY(hipGetLastError());
    if (timer) timer->mark(s, "factored_final");
    if (kind.acht) {
        const AchtSplit &as = L.as;
        cplx *part = reinterpret_cast<cplx *>(w + L.acht_part);
        cplx *part_occ = reinterpret_cast<cplx *>(w + L.acht_occ);
        if (p.nb > p.nv) {
            const unsigned n = (unsigned)((p.nb - p.nv + AchtU::TN - 1) / AchtU::TN) * as.jtiles * as.nks;
            hipLaunchKernelGGL((k_acht_f64<2, 4>), dim3(n), dim3(AchtU::NT), 0, s, ra.p2, p.aqsn, bg, p.idx,
                               p.vcoul, p.ngpown, p.ncouls, p.nb, p.nv, p.nb, as.chunk, as.nks, as.jtiles, part);
            GPP_TRY(hipGetLastError());
        }
        if (p.nv > 0 && !occ_in_reduce) {  // else summed by the reduce
            hipLaunchKernelGGL(k_acht_occ, dim3(p.ngpown, L.nsplit), dim3(kRedThreads), 0, s, ra.p2, ra.br1, p.aqsn,
                               bg, p.idx, p.vcoul, p.ngpown, p.ncouls, p.nv, kind.red_chunk, L.nsplit, part_occ);
            GPP_TRY(hipGetLastError());
        }
        hipLaunchKernelGGL(k_acht_final, dim3((p.nb + kAchtFinalBands - 1) / kAchtFinalBands + p.nv), dim3(256), 0,
                           s, part, as.jtiles * as.nks, part_occ,
                           p.ngpown * L.nsplit, p.nb, p.nv, res.acht_n1_loc());
        GPP_TRY(hipGetLastError());
        if (timer) timer->mark(s, "factored_acht");
    }
    return hipSuccess;
}

// ---------------------------------------------------------------------------
// Band-dependent frequencies.  The band sum factors over any set of bands
// that share one wx row, so bands are grouped by identical rows ("buckets")
// and each bucket is one factored sub-problem: its bands are gathered into a
// contiguous block (occupied bands first), the sub-problem runs through
// run_factored with that bucket's row, and the per-bucket results are summed
// in bucket order (acht_n1_loc scattered back to the original band order).

namespace {

__global__ void k_permute_rows(const cplx *__restrict__ src, const int32_t *__restrict__ perm, int nrows,
                               int ncols, cplx *__restrict__ dst)
{
    const size_t n = (size_t)nrows * ncols;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        const size_t r = t / ncols, c = t % ncols;
        dst[t] = src[(size_t)perm[r] * ncols + c];
    }
}

// One block.  tab[3b..3b+2] = (result offset, first permuted row, bands) of
// bucket b; bres holds each bucket's DevResults back to back.
__global__ void k_bucket_combine(const cplx *__restrict__ bres, const int32_t *__restrict__ tab,
                                 const int32_t *__restrict__ perm, int nbuckets, int nfreq, int nb,
                                 cplx *__restrict__ out)
{
    for (int iw = threadIdx.x; iw < 2 * nfreq; iw += blockDim.x) {
        cplx s = czero();
        for (int b = 0; b < nbuckets; ++b) s = cadd(s, bres[tab[3 * b] + iw]);
        out[iw] = s;  // achtemp[iw] for iw < nfreq, then asxtemp
    }
    if (threadIdx.x == 0) {
        cplx s = czero();
        for (int b = 0; b < nbuckets; ++b) s = cadd(s, bres[tab[3 * b] + 2 * nfreq + tab[3 * b + 2]]);
        out[2 * nfreq + nb] = s;  // achstemp
    }
    for (int b = 0; b < nbuckets; ++b)
        for (int k = threadIdx.x; k < tab[3 * b + 2]; k += blockDim.x)
            out[2 * nfreq + perm[tab[3 * b + 1] + k]] = bres[tab[3 * b] + 2 * nfreq + k];
}

struct BucketLayout {
    size_t aqsm, aqsn, perm, tab, bres, sub, total;
};

DevProblem bucket_problem(const DevProblem &p, const BucketPlan &bp, int b, const cplx *aqsm_p, const cplx *aqsn_p)
{
    DevProblem q = p;
    q.nb = bp.size[b];
    q.nv = bp.nocc[b];
    q.aqsm = aqsm_p + (size_t)bp.start[b] * p.ncouls;
    q.aqsn = aqsn_p + (size_t)bp.start[b] * p.ncouls;
    q.wx = p.wx + (size_t)bp.row[b] * p.nfreq;
    q.wx_stride = 0;
    q.wx_neg = bp.neg.empty() ? 1 : bp.neg[b];
    return q;
}

BucketLayout bucket_layout(const DevProblem &p, const GramKind &k, const BucketPlan &bp)
{
    BucketLayout L;
    const int nbk = (int)bp.size.size();
    size_t off = 0;
    L.aqsm = off; off += align256(sizeof(cplx) * (size_t)p.nb * p.ncouls);
    L.aqsn = off; off += align256(sizeof(cplx) * (size_t)p.nb * p.ncouls);
    L.perm = off; off += align256(sizeof(int32_t) * (size_t)p.nb);
    L.tab = off;  off += align256(sizeof(int32_t) * 3 * (size_t)nbk);
    size_t sub = 0;
    for (int b = 0; b < nbk; ++b) {
        const DevProblem q = bucket_problem(p, bp, b, nullptr, nullptr);
        sub = std::max(sub, factored_workspace_bytes(q, k));
    }
    L.bres = off; off += align256(sizeof(cplx) * bp.nres);
    L.sub = off;  off += sub;
    L.total = off;
    return L;
}

}  // namespace

BucketPlan plan_buckets(const double *wx_rows, int nb, int nv, int nfreq)
{
    BucketPlan bp;
    std::vector<int> bucket(nb, -1);
    for (int n1 = 0; n1 < nb; ++n1) {
        const double *r = wx_rows + (size_t)n1 * nfreq;
        for (int b = 0; b < (int)bp.row.size(); ++b)
            if (memcmp(r, wx_rows + (size_t)bp.row[b] * nfreq, sizeof(double) * nfreq) == 0) {
                bucket[n1] = b;
                break;
            }
        if (bucket[n1] < 0) {
            bucket[n1] = (int)bp.row.size();
            bp.row.push_back(n1);
            bp.neg.push_back(wx_has_negative(r, nfreq));
        }
    }
    const int nbk = (int)bp.row.size();
    bp.start.assign(nbk, 0);
    bp.size.assign(nbk, 0);
    bp.nocc.assign(nbk, 0);
    bp.tab.assign(3 * (size_t)nbk, 0);
    bp.nres = 0;
    for (int b = 0; b < nbk; ++b) {
        bp.start[b] = (int)bp.perm.size();
        for (int pass = 0; pass < 2; ++pass)  // occupied bands first, each in band order
            for (int n1 = 0; n1 < nb; ++n1)
                if (bucket[n1] == b && (n1 < nv) == (pass == 0)) {
                    bp.perm.push_back(n1);
                    ++bp.size[b];
                    if (pass == 0) ++bp.nocc[b];
                }
        bp.tab[3 * b] = (int32_t)bp.nres;
        bp.tab[3 * b + 1] = bp.start[b];
        bp.tab[3 * b + 2] = bp.size[b];
        bp.nres += (size_t)2 * nfreq + bp.size[b] + 1;
    }
    return bp;
}

size_t bucketed_workspace_bytes(const DevProblem &p, GramKind kind, const BucketPlan &bp)
{
    return bucket_layout(p, kind, bp).total;
}

hipError_t run_factored_bucketed(const DevProblem &p, GramKind kind, const BucketPlan &bp, void *work,
                                 const DevResults &res, hipStream_t s, StageTimer *timer)
{
    if (p.wx_stride != p.nfreq || (int)bp.perm.size() != p.nb || bp.tab.size() != 3 * bp.size.size())
        return hipErrorInvalidValue;
    const BucketLayout L = bucket_layout(p, kind, bp);
    char *w = static_cast<char *>(work);
    cplx *aqsm_p = reinterpret_cast<cplx *>(w + L.aqsm);
    cplx *aqsn_p = reinterpret_cast<cplx *>(w + L.aqsn);
    int32_t *perm = reinterpret_cast<int32_t *>(w + L.perm);
    int32_t *tab = reinterpret_cast<int32_t *>(w + L.tab);
    cplx *bres = reinterpret_cast<cplx *>(w + L.bres);
    const int nbk = (int)bp.size.size();

    if (timer) timer->begin(s);
    GPP_TRY(hipMemcpyAsync(perm, bp.perm.data(), sizeof(int32_t) * p.nb, hipMemcpyHostToDevice, s));
    GPP_TRY(hipMemcpyAsync(tab, bp.tab.data(), sizeof(int32_t) * 3 * nbk, hipMemcpyHostToDevice, s));
    {
        const size_t n = (size_t)p.nb * p.ncouls;
        size_t blocks = (n + 255) / 256;
        if (blocks > 8192) blocks = 8192;
        hipLaunchKernelGGL(k_permute_rows, dim3((unsigned)blocks), dim3(256), 0, s, p.aqsm, perm, p.nb, p.ncouls,
                           aqsm_p);
        hipLaunchKernelGGL(k_permute_rows, dim3((unsigned)blocks), dim3(256), 0, s, p.aqsn, perm, p.nb, p.ncouls,
                           aqsn_p);
        GPP_TRY(hipGetLastError());
    }
    if (timer) timer->mark(s, "bucket_permute");
    for (int b = 0; b < nbk; ++b) {
        const DevProblem q = bucket_problem(p, bp, b, aqsm_p, aqsn_p);
        DevResults r;
        r.base = bres + bp.tab[3 * b];
        r.nfreq = p.nfreq;
        r.nb = q.nb;
        GPP_TRY(run_factored(q, kind, w + L.sub, r, s, nullptr));
    }
    if (timer) timer->mark(s, "bucket_factored");
    hipLaunchKernelGGL(k_bucket_combine, dim3(1), dim3(256), 0, s, bres, tab, perm, nbk, p.nfreq, p.nb, res.base);
    GPP_TRY(hipGetLastError());
    if (timer) timer->mark(s, "bucket_combine");
    return hipSuccess;
}

}  // namespace gpp
