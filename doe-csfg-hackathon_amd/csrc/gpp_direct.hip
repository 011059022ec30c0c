// gpp_direct.hip -- direct FP64 form of the GPP sigma band loop on gfx950.
//
// Reference: gppComplex.cpp:335-383 (band loop), :131-150 (noflagOCC_solver,
// unoccupied bands), :107-128 + :16-63 (flagOCC_solver + ssxt_scht_solver,
// occupied bands), :66-101 (reduce_achstemp).
//
// Every update (n1, j, ig, iw) evaluates its own pole factor, the complex
// division included; only the frequency-independent matrix element
// a = conj(aqsm[n1,igp]) * aqsn[n1,ig] is shared across the frequencies of
// one (n1, j, ig), which the reference also computes identically per iw.
//
// Work decomposition (unoccupied bands, the hot loop):
//   block  = NW waves; block (x, y) owns bands nv + x*NW .. +NW-1 (one band
//            per wave) and the G'-column tile j0 = y*BJ .. +BJ-1;
//   lane   = one G-vector ig of a 64-wide chunk; the chunk loop streams all
//            ncouls G-vectors;
//   LDS    = the chunk's wtilde / I_eps rows of the BJ columns, loaded once
//            per block and read by all NW waves (double-buffered, one barrier
//            per chunk);
//   VGPRs  = the BJ x FC per-frequency complex accumulators of each lane.
// After the chunk loop each wave reduces its accumulators by xor butterfly,
// scales by 0.5*vcoul[igp] and sums its BJ columns in column order, writing
// one partial per (band, column tile, frequency).  Two fixed-order tree
// passes then reduce the partials, so results are bitwise reproducible.
#include "gpp_internal.hpp"

namespace gpp {
namespace {

constexpr int kBJ = 4;   // G' columns per block (unoccupied kernel tile)
constexpr int kNW = 8;   // waves (bands) per block, unoccupied kernel
constexpr int kOccWaves = 4;
constexpr int kOccCols = kBJ / kOccWaves;  // columns per wave, occupied kernel
constexpr int kReduceThreads = 256;

struct DirectArgs {
    const cplx *aqsm, *aqsn, *ieps, *wtilde;
    const int32_t *idx;
    const double *vcoul, *wx;
    int nb, nv, ncouls, ngpown, nfreq, njt;
    int wx_stride;  // 0: shared wx row; nfreq: band n1 reads wx[n1 * nfreq + iw]
    double occ;
    cplx *part_ach;   // [nb][njt][nfreq]
    cplx *part_asx;   // [max(nv,1)][njt][nfreq]
    cplx *part_achs;  // [nb][kAchSplit], then the AchCol entries (ach_cols_entry)
    cplx *band_ach;   // [nb][nfreq]
    cplx *band_asx;   // [max(nv,1)][nfreq]
    cplx *bg;         // optional gathered rows of aqsm (factored forms), else null
};

struct WorkLayout {
    size_t part_ach, part_asx, part_achs, band_ach, band_asx, total;
};

__host__ inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

WorkLayout layout_for(const DevProblem &p)
{
    const size_t njt = (size_t)(p.ngpown + kBJ - 1) / kBJ;
    const size_t nvx = p.nv > 0 ? (size_t)p.nv : 1;
    WorkLayout L;
    size_t off = 0;
    L.part_ach = off;  off += align256(sizeof(cplx) * (size_t)p.nb * njt * p.nfreq + sizeof(cplx));
    L.part_asx = off;  off += align256(sizeof(cplx) * nvx * njt * p.nfreq);
    L.part_achs = off; off += align256(sizeof(cplx) * ach_region_entries(p.nb, p.ngpown));
    L.band_ach = off;  off += align256(sizeof(cplx) * ((size_t)p.nb * p.nfreq + 1));
    L.band_asx = off;  off += align256(sizeof(cplx) * nvx * p.nfreq);
    L.total = off;
    return L;
}

DirectArgs make_args(const DevProblem &p, void *work)
{
    WorkLayout L = layout_for(p);
    char *w = static_cast<char *>(work);
    DirectArgs a;
    a.aqsm = p.aqsm; a.aqsn = p.aqsn; a.ieps = p.ieps; a.wtilde = p.wtilde;
    a.idx = p.idx; a.vcoul = p.vcoul; a.wx = p.wx; a.wx_stride = p.wx_stride;
    a.nb = p.nb; a.nv = p.nv; a.ncouls = p.ncouls; a.ngpown = p.ngpown; a.nfreq = p.nfreq;
    a.njt = (p.ngpown + kBJ - 1) / kBJ;
    a.occ = p.occ;
    a.part_ach = reinterpret_cast<cplx *>(w + L.part_ach);
    a.part_asx = reinterpret_cast<cplx *>(w + L.part_asx);
    a.part_achs = reinterpret_cast<cplx *>(w + L.part_achs);
    a.band_ach = reinterpret_cast<cplx *>(w + L.band_ach);
    a.band_asx = reinterpret_cast<cplx *>(w + L.band_asx);
    a.bg = nullptr;
    return a;
}

// ---------------------------------------------------------------------------
// Per-update arithmetic of the unoccupied bands.
//
// Fast form: with a = conj(aqsm)*aqsn, the reference's term
//   a * ((wtilde * conj(wdiff) * rden) * I_eps),  rden = 1/|wdiff|^2
// is evaluated as ((a * (wtilde * I_eps)) * conj(wdiff)) * rden, the same
// product re-associated; wdiff = wx - wtilde and the reciprocal are formed
// per update as in gppComplex.cpp:139-143, the reciprocal refined to third
// order (rcp_2o: about one rounding, unbiased, as the reference's quotient).
// wtie = wtilde * I_eps, formed once per (column, G-vector) when staged.
template <int FC>
__device__ __forceinline__ void unocc_update_fast(cplx (&acc)[FC], cplx cm, cplx an, cplx wt,
                                                  cplx wtie, const double (&wxv)[FC])
{
    const cplx a = cmul_fast(cm, an);
    const cplx q = cmul_fast(a, wtie);
#pragma unroll
    for (int f = 0; f < FC; ++f) {
        const double wdr = wxv[f] - wt.re;   // wdiff = (wx - wt.re, -wt.im)
        const double d = wdr * wdr + wt.im * wt.im;
        const double r = rcp_2o(d);
        // q * conj(wdiff) = q * (wdr, wt.im)
        const double tr = q.re * wdr - q.im * wt.im;
        const double ti = q.im * wdr + q.re * wt.im;
        acc[f].re += tr * r;
        acc[f].im += ti * r;
    }
}

// Strict form: the reference's operation order for one term,
// gppComplex.cpp:139-146, IEEE division, no FMA contraction.
template <int FC>
__device__ __forceinline__ void unocc_update_strict(cplx (&acc)[FC], cplx cm, cplx an, cplx wt,
                                                    cplx ie, const double (&wxv)[FC])
{
#pragma clang fp contract(off)
    const cplx a = cmul(cm, an);  // product(mygpvar1, aqsn), :146
#pragma unroll
    for (int f = 0; f < FC; ++f) {
        const cplx wdiff = cdminus(wxv[f], wt);                 // :139
        const double wdiffr = cmul(wdiff, cconj(wdiff)).re;     // :140
        const double rden = 1 / wdiffr;                         // :141
        const cplx delw = cscale(cmul(wt, cconj(wdiff)), rden); // :143
        cacc(acc[f], cmul(a, cmul(delw, ie)));                  // :146
    }
}

// ---------------------------------------------------------------------------
// Unoccupied bands.  Grid (ceil((nb-nv)/NW), njt, ceil(nfreq/FC)), block
// NW*64 threads.
template <int FC, bool STRICT>
__global__ __launch_bounds__(kNW *kWave) void k_unocc(DirectArgs a)
{
    constexpr int TILE = kBJ * kWave;  // (column, G-vector) positions per chunk
    static_assert(TILE <= kNW * kWave, "one position per staging thread");
    __shared__ cplx s_wt[2][kBJ][kWave];
    __shared__ cplx s_ie[2][kBJ][kWave];  // I_eps (strict) or wtilde * I_eps (fast)

    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int lane = threadIdx.x % kWave;
    const int n1 = a.nv + blockIdx.x * kNW + wave;
    const bool band_ok = n1 < a.nb;
    const int jt = blockIdx.y;
    const int j0 = jt * kBJ;
    const int iw0 = blockIdx.z * FC;
    const int ncouls = a.ncouls;

    // conj(aqsm[n1, igp(j)]) per column (mygpvar1, gppComplex.cpp:134).
    cplx cm[kBJ];
#pragma unroll
    for (int jj = 0; jj < kBJ; ++jj) {
        const int j = j0 + jj;
        cm[jj] = czero();
        if (band_ok && j < a.ngpown) cm[jj] = cconj(a.aqsm[(size_t)n1 * ncouls + igp_of(a.idx, j, ncouls)]);
    }
    double wxv[FC];
#pragma unroll
    for (int f = 0; f < FC; ++f)
        wxv[f] = (band_ok && iw0 + f < a.nfreq) ? a.wx[(size_t)n1 * a.wx_stride + iw0 + f] : 0.0;

    cplx acc[kBJ][FC];
#pragma unroll
    for (int jj = 0; jj < kBJ; ++jj)
#pragma unroll
        for (int f = 0; f < FC; ++f) acc[jj][f] = czero();

    // Staging of one chunk's wtilde/I_eps tile: thread t < TILE loads both at
    // position t; the fast form stores wtilde * I_eps, formed here once
    // instead of once per band.  Out-of-range entries get wtilde = i (so
    // |wdiff|^2 >= 1) and I_eps = 0: their terms are exact zeros.
    const bool stager = threadIdx.x < TILE;
    const int s_jj = (threadIdx.x % TILE) / kWave, s_l = threadIdx.x % kWave;
    cplx st_wt = czero(), st_ie = czero();
    auto load_chunk = [&](int c) {
        if (!stager) return;
        const int j = j0 + s_jj, ig = c * kWave + s_l;
        const bool ok = j < a.ngpown && ig < ncouls;
        st_wt = ok ? a.wtilde[(size_t)j * ncouls + ig] : mk(0.0, 1.0);
        st_ie = ok ? a.ieps[(size_t)j * ncouls + ig] : czero();
    };
    auto store_chunk = [&](int buf) {
        if (!stager) return;
        s_wt[buf][s_jj][s_l] = st_wt;
        s_ie[buf][s_jj][s_l] = STRICT ? st_ie : cmul_fast(st_wt, st_ie);
    };

    const int nchunks = (ncouls + kWave - 1) / kWave;
    load_chunk(0);
    store_chunk(0);
    __syncthreads();
    for (int c = 0; c < nchunks; ++c) {
        const int buf = c & 1;
        if (c + 1 < nchunks) load_chunk(c + 1);
        if (band_ok) {
            const int ig = c * kWave + lane;
            const cplx an = ig < ncouls ? a.aqsn[(size_t)n1 * ncouls + ig] : czero();
#pragma unroll
            for (int jj = 0; jj < kBJ; ++jj) {
                const cplx wt = s_wt[buf][jj][lane];
                const cplx ie = s_ie[buf][jj][lane];
                if (STRICT) unocc_update_strict<FC>(acc[jj], cm[jj], an, wt, ie, wxv);
                else unocc_update_fast<FC>(acc[jj], cm[jj], an, wt, ie, wxv);
            }
        }
        if (c + 1 < nchunks) store_chunk(buf ^ 1);
        __syncthreads();
    }

    if (!band_ok) return;
    // scht per (column, frequency); then sum_j (0.5*scht)*vcoul[igp] in
    // column order (gppComplex.cpp:374, :379).
#pragma unroll
    for (int f = 0; f < FC; ++f) {
        cplx s = czero();
#pragma unroll
        for (int jj = 0; jj < kBJ; ++jj) {
            const cplx scht = wave_sum(acc[jj][f]);
            const int j = j0 + jj;
            if (j < a.ngpown) cacc(s, cscale(cscale(scht, 0.5), a.vcoul[igp_of(a.idx, j, ncouls)]));
        }
        if (lane == 0 && iw0 + f < a.nfreq) a.part_ach[((size_t)n1 * a.njt + jt) * a.nfreq + iw0 + f] = s;
    }
}

// ---------------------------------------------------------------------------
// Occupied bands n1 < nv.  Grid (njt, nv, ceil(nfreq/FC)), block 4 waves;
// wave w owns columns j0 + w*kOccCols .. +kOccCols-1 of the block's tile and
// strides over ig itself.  The four waves' column sums are combined in
// column order through LDS.
template <int FC>
__global__ __launch_bounds__(kOccWaves *kWave) void k_occ(DirectArgs a)
{
    __shared__ cplx s_ach[kBJ][FC];
    __shared__ cplx s_asx[kBJ][FC];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int lane = threadIdx.x % kWave;
    const int jt = blockIdx.x;
    const int n1 = blockIdx.y;
    const int iw0 = blockIdx.z * FC;
    const int ncouls = a.ncouls;

    double wxv[FC], wx2[FC], rwx4[FC];
#pragma unroll
    for (int f = 0; f < FC; ++f) {
        wxv[f] = (iw0 + f < a.nfreq) ? a.wx[(size_t)n1 * a.wx_stride + iw0 + f] : 1.0;
        // std::pow(wxt,2) and std::pow(cden,2), gppComplex.cpp:38-40
        wx2[f] = wxv[f] * wxv[f];
        rwx4[f] = 1.00 / (wx2[f] * wx2[f]);
    }

    cplx ssxt[kOccCols][FC], scht[kOccCols][FC];
#pragma unroll
    for (int c = 0; c < kOccCols; ++c)
#pragma unroll
        for (int f = 0; f < FC; ++f) ssxt[c][f] = scht[c][f] = czero();

    int jcol[kOccCols];
    cplx mv1[kOccCols];
#pragma unroll
    for (int c = 0; c < kOccCols; ++c) {
        jcol[c] = jt * kBJ + wave * kOccCols + c;
        mv1[c] = czero();
        if (jcol[c] < a.ngpown) mv1[c] = cconj(a.aqsm[(size_t)n1 * ncouls + igp_of(a.idx, jcol[c], ncouls)]);
    }

    for (int ig = lane; ig < ncouls; ig += kWave) {
        const cplx an = a.aqsn[(size_t)n1 * ncouls + ig];
#pragma unroll
        for (int c = 0; c < kOccCols; ++c) {
            if (jcol[c] >= a.ngpown) continue;
            const size_t off = (size_t)jcol[c] * ncouls + ig;
            const cplx wt = a.wtilde[off];
            const cplx ie = a.ieps[off];
            const cplx m = cmul(an, mv1[c]);  // matngmatmgp, gppComplex.cpp:121
#pragma unroll
            for (int f = 0; f < FC; ++f) {
                cplx sch, ssx;
                occ_pole_terms(wxv[f], wx2[f], rwx4[f], wt, ie, sch, ssx);
                cacc(ssxt[c][f], cmul(m, ssx));  // :61, :125
                cacc(scht[c][f], cmul(m, sch));  // :62, :126
            }
        }
    }

#pragma unroll
    for (int c = 0; c < kOccCols; ++c) {
        const int jj = wave * kOccCols + c;
        const bool ok = jcol[c] < a.ngpown;
        const double vc = ok ? a.vcoul[igp_of(a.idx, jcol[c], ncouls)] : 0.0;
#pragma unroll
        for (int f = 0; f < FC; ++f) {
            const cplx sx = wave_sum(ssxt[c][f]);
            const cplx sc = wave_sum(scht[c][f]);
            if (lane == 0) {
                // achtemp: (0.5*scht)*vcoul (:361, :379); asxtemp: ssxt*(occ*vcoul) (:363)
                s_ach[jj][f] = ok ? cscale(cscale(sc, 0.5), vc) : czero();
                s_asx[jj][f] = ok ? cscale(sx, a.occ * vc) : czero();
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < FC && iw0 + (int)threadIdx.x < a.nfreq) {
        const int f = threadIdx.x;
        cplx sa = czero(), sx = czero();
        for (int jj = 0; jj < kBJ; ++jj) {
            cacc(sa, s_ach[jj][f]);
            cacc(sx, s_asx[jj][f]);
        }
        const size_t o = ((size_t)n1 * a.njt + jt) * a.nfreq + iw0 + f;
        a.part_ach[o] = sa;
        a.part_asx[o] = sx;
    }
}

// ---------------------------------------------------------------------------
// achstemp per band, reduce_achstemp (gppComplex.cpp:66-101): one wave per
// band and block, so that the few hundred bands spread over every CU.
constexpr int kAchWaves = 1;
__global__ __launch_bounds__(kAchWaves * kWave) void k_achstemp(DirectArgs a)
{
    const int item = blockIdx.x * kAchWaves + threadIdx.x / kWave;  // band * kAchSplit + column group
    if (item >= a.nb * kAchSplit) return;
    const AchCol *cols = reinterpret_cast<const AchCol *>(a.part_achs + ach_cols_entry(a.nb));
    const cplx s = achstemp_band(a.aqsm, a.aqsn, cols, a.ncouls, a.ngpown, item / kAchSplit, item % kAchSplit, a.bg);
    if (threadIdx.x % kWave == 0) a.part_achs[item] = s;
}

__global__ __launch_bounds__(256) void k_ach_cols(DirectArgs a)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < a.ngpown)
        ach_col(a.ieps, a.idx, a.vcoul, a.ncouls, j, reinterpret_cast<AchCol *>(a.part_achs + ach_cols_entry(a.nb)));
}

// The columns' achstemp entries, then the (band, column group) partials (and
// the Bg gather when a.bg is set).
hipError_t launch_ach(const DirectArgs &a, hipStream_t s)
{
    if (a.ngpown > 0) hipLaunchKernelGGL(k_ach_cols, dim3((a.ngpown + 255) / 256), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_achstemp, dim3((a.nb * kAchSplit + kAchWaves - 1) / kAchWaves), dim3(kAchWaves * kWave), 0, s,
                       a);
    return hipGetLastError();
}

// Per band: partials over column tiles -> band_ach[n1][iw] (and band_asx
// for occupied bands).  Grid nb, block kReduceThreads.
__global__ __launch_bounds__(kReduceThreads) void k_reduce_bands(DirectArgs a, int n1_begin)
{
    __shared__ cplx scratch[kReduceThreads / kWave];
    const int n1 = n1_begin + blockIdx.x;
    for (int iw = 0; iw < a.nfreq; ++iw) {
        cplx s = czero(), x = czero();
        for (int t = threadIdx.x; t < a.njt; t += blockDim.x) {
            const size_t o = ((size_t)n1 * a.njt + t) * a.nfreq + iw;
            s = cadd(s, a.part_ach[o]);
            if (n1 < a.nv) x = cadd(x, a.part_asx[o]);
        }
        s = block_sum(s, scratch);
        if (threadIdx.x == 0) a.band_ach[(size_t)n1 * a.nfreq + iw] = s;
        if (n1 < a.nv) {
            x = block_sum(x, scratch);
            if (threadIdx.x == 0) a.band_asx[(size_t)n1 * a.nfreq + iw] = x;
        }
    }
}

// Final pass, one block: achtemp, asxtemp, acht_n1_loc, achstemp for bands
// [n1_begin, n1_end) (achstemp only when with_achs).
__global__ __launch_bounds__(kReduceThreads) void k_reduce_final(DirectArgs a, cplx *out, int n1_begin,
                                                                 int n1_end, int with_achs)
{
    __shared__ cplx scratch[kReduceThreads / kWave];
    const int nf = a.nfreq;
    const int acht_iw = nf > 2 ? 2 : nf - 1;  // gppComplex.cpp:381
    for (int iw = 0; iw < nf; ++iw) {
        cplx s = czero(), x = czero();
        for (int n1 = n1_begin + threadIdx.x; n1 < n1_end; n1 += blockDim.x) {
            s = cadd(s, a.band_ach[(size_t)n1 * nf + iw]);
            if (n1 < a.nv) x = cadd(x, a.band_asx[(size_t)n1 * nf + iw]);
        }
        s = block_sum(s, scratch);
        x = block_sum(x, scratch);
        if (threadIdx.x == 0) {
            out[iw] = s;
            out[nf + iw] = x;
        }
    }
    for (int n1 = threadIdx.x; n1 < a.nb; n1 += blockDim.x)
        out[2 * nf + n1] = (n1 >= n1_begin && n1 < n1_end) ? a.band_ach[(size_t)n1 * nf + acht_iw] : czero();
    cplx h = czero();
    if (with_achs)
        for (int t = threadIdx.x; t < a.nb * kAchSplit; t += blockDim.x) h = cadd(h, a.part_achs[t]);
    h = block_sum(h, scratch);
    if (threadIdx.x == 0) out[2 * nf + a.nb] = h;
}

__global__ void k_zero_ints(int *p, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

__global__ void k_check_index(const int32_t *idx, int n, int *count)
{
    int bad = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) bad += idx[i] < 0;
    if (bad) atomicAdd(count, bad);
}

// Accumulates results into the reference launchers' split outputs.
// which: 0 occupied (achtemp, asxtemp, acht_n1_loc[0..nv)), 1 unoccupied
// (achtemp, acht_n1_loc[nv..nb)), 2 achstemp.
__global__ void k_accumulate_compat(const cplx *res, int nfreq, int nb, int which, int nv,
                                    double *ach_re, double *ach_im, cplx *asx, cplx *acht_n1,
                                    double *achs_re, double *achs_im)
{
    const int t = threadIdx.x;
    if (which == 2) {
        if (t == 0) {
            achs_re[0] += res[2 * nfreq + nb].re;
            achs_im[0] += res[2 * nfreq + nb].im;
        }
        return;
    }
    if (t < nfreq) {
        ach_re[t] += res[t].re;
        ach_im[t] += res[t].im;
        if (which == 0 && asx) cacc(asx[t], res[nfreq + t]);
    }
    if (acht_n1) {
        const int lo = which == 0 ? 0 : nv, hi = which == 0 ? nv : nb;
        for (int n1 = lo + t; n1 < hi; n1 += blockDim.x) cacc(acht_n1[n1], res[2 * nfreq + n1]);
    }
}

template <int FC>
hipError_t launch_unocc_fc(const DirectArgs &a, bool strict, hipStream_t s)
{
    const int nun = a.nb - a.nv;
    if (nun <= 0 || a.ngpown <= 0) return hipSuccess;
    dim3 grid((nun + kNW - 1) / kNW, a.njt, (a.nfreq + FC - 1) / FC);
    if (strict) hipLaunchKernelGGL((k_unocc<FC, true>), grid, dim3(kNW * kWave), 0, s, a);
    else hipLaunchKernelGGL((k_unocc<FC, false>), grid, dim3(kNW * kWave), 0, s, a);
    return hipGetLastError();
}

template <int FC>
hipError_t launch_occ_fc(const DirectArgs &a, hipStream_t s)
{
    if (a.nv <= 0 || a.ngpown <= 0) return hipSuccess;
    dim3 grid(a.njt, a.nv, (a.nfreq + FC - 1) / FC);
    hipLaunchKernelGGL((k_occ<FC>), grid, dim3(kOccWaves * kWave), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_unocc(const DirectArgs &a, bool strict, hipStream_t s)
{
    switch (a.nfreq) {
    case 1: return launch_unocc_fc<1>(a, strict, s);
    case 2: return launch_unocc_fc<2>(a, strict, s);
    case 3: return launch_unocc_fc<3>(a, strict, s);
    default: return launch_unocc_fc<4>(a, strict, s);
    }
}

hipError_t launch_occ(const DirectArgs &a, hipStream_t s)
{
    switch (a.nfreq) {
    case 1: return launch_occ_fc<1>(a, s);
    case 2: return launch_occ_fc<2>(a, s);
    case 3: return launch_occ_fc<3>(a, s);
    default: return launch_occ_fc<4>(a, s);
    }
}

#define GPP_TRY(x)                         \
    do {                                   \
        hipError_t e_ = (x);               \
        if (e_ != hipSuccess) return e_;   \
    } while (0)

}  // namespace

size_t direct_workspace_bytes(const DevProblem &p) { return layout_for(p).total; }

hipError_t run_direct(const DevProblem &p, bool strict, void *work, const DevResults &res,
                      hipStream_t s, StageTimer *timer)
{
    const DirectArgs a = make_args(p, work);
    if (timer) timer->begin(s);
    GPP_TRY(launch_occ(a, s));
    if (timer) timer->mark(s, "direct_occupied");
    GPP_TRY(launch_unocc(a, strict, s));
    if (timer) timer->mark(s, strict ? "direct_unoccupied_strict" : "direct_unoccupied");
    if (p.nb > 0) {
        GPP_TRY(launch_ach(a, s));
        hipLaunchKernelGGL(k_reduce_bands, dim3(p.nb), dim3(kReduceThreads), 0, s, a, 0);
        GPP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kReduceThreads), 0, s, a, res.base, 0, p.nb, 1);
    GPP_TRY(hipGetLastError());
    if (timer) timer->mark(s, "direct_reduce");
    return hipSuccess;
}

hipError_t run_direct_part(const DevProblem &p, int which, void *work, const DevResults &res,
                           hipStream_t s)
{
    const DirectArgs a = make_args(p, work);
    if (which == 0) {
        GPP_TRY(launch_occ(a, s));
        if (p.nv > 0) {
            hipLaunchKernelGGL(k_reduce_bands, dim3(p.nv), dim3(kReduceThreads), 0, s, a, 0);
            GPP_TRY(hipGetLastError());
        }
        hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kReduceThreads), 0, s, a, res.base, 0, p.nv, 0);
    } else if (which == 1) {
        GPP_TRY(launch_unocc(a, false, s));
        if (p.nb > p.nv) {
            hipLaunchKernelGGL(k_reduce_bands, dim3(p.nb - p.nv), dim3(kReduceThreads), 0, s, a, p.nv);
            GPP_TRY(hipGetLastError());
        }
        hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kReduceThreads), 0, s, a, res.base, p.nv, p.nb, 0);
    } else {
        if (p.nb > 0) GPP_TRY(launch_ach(a, s));
        hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kReduceThreads), 0, s, a, res.base, 0, 0, 1);
    }
    return hipGetLastError();
}

hipError_t launch_achstemp(const DevProblem &p, cplx *part_achs, cplx *bg, hipStream_t s)
{
    if (p.nb <= 0) return hipSuccess;
    DirectArgs a = make_args(p, nullptr);
    a.part_achs = part_achs;
    a.bg = bg;
    return launch_ach(a, s);
}

hipError_t launch_check_index(const int32_t *d_idx, int ngpown, int *d_count, hipStream_t s)
{
    if (ngpown <= 0) return hipSuccess;
    const int blocks = (ngpown + 255) / 256 < 1024 ? (ngpown + 255) / 256 : 1024;
    hipLaunchKernelGGL(k_check_index, dim3(blocks), dim3(256), 0, s, d_idx, ngpown, d_count);
    return hipGetLastError();
}

hipError_t launch_zero_ints(int *d_p, int n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_ints, dim3((n + 255) / 256 < 64 ? (n + 255) / 256 : 64), dim3(256), 0, s, d_p, n);
    return hipGetLastError();
}

hipError_t launch_accumulate_compat(const DevResults &res, int which, int nv, double *achtemp_re,
                                    double *achtemp_im, cplx *asxtemp, cplx *acht_n1_loc,
                                    double *achstemp_re, double *achstemp_im, hipStream_t s)
{
    hipLaunchKernelGGL(k_accumulate_compat, dim3(1), dim3(256), 0, s, res.base, res.nfreq, res.nb, which,
                       nv, achtemp_re, achtemp_im, asxtemp, acht_n1_loc, achstemp_re, achstemp_im);
    return hipGetLastError();
}

}  // namespace gpp
