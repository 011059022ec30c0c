// gpp_batch.hip -- many independent GPP sigma problems of uneven size in one
// pass of the factored FP64 form.
//
// Each stage of the factored form (gpp_factored_dev.hpp) becomes ONE launch
// over a work list that spans every instance (the reduce: one per occupied
// band class):
//   achstemp  (instance, band) items, one wave each, also gathering Bg
//   Gram      (instance, G' tile, G tile) items of the FP64-MFMA product
//   reduce    (instance, column, G-vector split) items, every frequency
//   final     one block per instance
// Lists are built on the host, sorted by decreasing cost; the persistent
// launches walk them with a static round-robin (block b takes items b, b+G,
// b+2G, ...), the reduce runs one block per item.  Every item is computed by
// exactly one block (or wave), so results do not depend on scheduling.  The
// launches are captured once into a hipGraph and replayed.
//
// The Ozaki kinds replace the Gram launch by three: column exponents and
// digit planes of every instance's Bg and aqsn (gpp_ozaki_dev.hpp), then the
// layout-3 INT8 / FP16 digit product over (instance, 64 x 64 tile) items.
#include <algorithm>
#include <vector>

#include "gpp_factored_dev.hpp"
#include "gpp_ozaki_dev.hpp"

namespace gpp {
namespace {

using namespace fact;

struct BInst {
    const cplx *aqsm, *aqsn, *ieps, *wtilde;
    const int32_t *idx;
    const double *vcoul, *wx;
    int nb, nv, ncouls, ngpown, nfreq, nsplit;
    double occ;
    cplx *bg, *mu, *col_ach, *col_asx, *part_achs, *out;
    // Ozaki kinds: digit planes and column exponents of Bg (b) and aqsn (a)
    void *oz_bd, *oz_ad;
    int *oz_eb, *oz_ea;
    int oz_ngp_pad, oz_nc_pad, oz_kpad;
};

// The columns' achstemp entries: grid (column blocks of 256, instance).
__global__ __launch_bounds__(256) void k_batch_ach_cols(const BInst *__restrict__ insts)
{
    const BInst &b = insts[blockIdx.y];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < b.ngpown)
        ach_col(b.ieps, b.idx, b.vcoul, b.ncouls, j, reinterpret_cast<AchCol *>(b.part_achs + ach_cols_entry(b.nb)));
}

// One wave per (instance, band, column group) item.
__global__ __launch_bounds__(kRedThreads) void k_batch_achstemp(const BInst *__restrict__ insts,
                                                                const int2 *__restrict__ items, int nitems)
{
    constexpr int W = kRedThreads / kWave;
    for (int w = blockIdx.x * W + threadIdx.x / kWave; w < nitems; w += gridDim.x * W) {
        const int2 it = items[w];
        const BInst &b = insts[it.x];  // it.y = band * kAchSplit + column group
        const AchCol *cols = reinterpret_cast<const AchCol *>(b.part_achs + ach_cols_entry(b.nb));
        const cplx s = achstemp_band(b.aqsm, b.aqsn, cols, b.ncouls, b.ngpown, it.y / kAchSplit, it.y % kAchSplit,
                                     b.bg);
        if (threadIdx.x % kWave == 0) b.part_achs[it.y] = s;
    }
}

template <int WJN, int WIN>
__global__ __launch_bounds__(64 * WJN * WIN, WJN * WIN >= 4 ? 8 / (WJN * WIN) : 2) void k_batch_gram(const BInst *__restrict__ insts,
                                                               const GramItem *__restrict__ items, int nitems)
{
    typedef GramCfg<WJN, WIN> C;
    __shared__ cplx sB[2][C::KS][C::TJ];
    __shared__ cplx sA[2][C::KS][C::TI];
    for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
        const GramItem it = items[w];
        const BInst &b = insts[it.inst];
        gram_f64_tile<WJN, WIN>(b.bg, b.aqsn, b.ngpown, b.ncouls, b.nv, b.nb, it.j0, it.i0, sB, sA, b.mu);
    }
}

// Column exponents of item (inst, side, 64-column block): side 0 = Bg's
// columns over bands [nv, nb), side 1 = aqsn's.
__global__ __launch_bounds__(256) void k_batch_oz_colexp(const BInst *__restrict__ insts,
                                                         const int4 *__restrict__ items, int nitems)
{
    __shared__ double part[4][64];
    for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
        const int4 it = items[w];
        const BInst &b = insts[it.x];
        if (it.y == 0)
            oz::colexp_block(b.bg, b.ngpown, b.nv, b.nb, b.ngpown, b.oz_ngp_pad, b.oz_eb, it.z, part);
        else
            oz::colexp_block(b.aqsn, b.ncouls, b.nv, b.nb, b.ncouls, b.oz_nc_pad, b.oz_ea, it.z, part);
        __syncthreads();  // part is reused by the next item
    }
}

// Digit pieces of item (inst, side, 256-column block, piece).
template <int S, typename D>
__global__ __launch_bounds__(256) void k_batch_oz_split(const BInst *__restrict__ insts,
                                                        const int4 *__restrict__ items, int nitems)
{
    for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
        const int4 it = items[w];
        const BInst &b = insts[it.x];
        const int K = b.nb - b.nv;
        const int c = it.z * 256 + threadIdx.x;
        if (it.y == 0) {
            if (c < b.oz_ngp_pad)
                oz::split_piece<S, D>(b.bg, b.ngpown, b.nv, K, b.ngpown, b.oz_ngp_pad, b.oz_kpad, b.oz_eb,
                                      static_cast<D *>(b.oz_bd), c, it.w);
        } else if (c < b.oz_nc_pad) {
            oz::split_piece<S, D>(b.aqsn, b.ncouls, b.nv, K, b.ncouls, b.oz_nc_pad, b.oz_kpad, b.oz_ea,
                                  static_cast<D *>(b.oz_ad), c, it.w);
        }
    }
}

// The (instance, 64 x 64 tile) items of a persistent block, items b, b + G,
// b + 2G, ... (G = grid size).
struct OzItemSeq {
    const BInst *insts;
    const GramItem *items;
    int nitems;
    __device__ bool get(int k, oz::OzTile &t) const
    {
        const long long w = (long long)blockIdx.x + (long long)k * gridDim.x;
        if (w >= nitems) return false;
        const GramItem it = items[w];
        const BInst &b = insts[it.inst];
        t = oz::OzTile{b.oz_bd, b.oz_ad, b.oz_eb, b.oz_ea, b.ngpown, b.ncouls, b.oz_ngp_pad, b.oz_nc_pad,
                       b.oz_kpad, it.j0, it.i0, b.mu};
        return true;
    }
};

// Layout-3 digit-product tiles over all instances as one stream of stages per
// persistent block (the next item's first stages load while one finishes).
template <int S, typename D, int NBUF, bool LONG>
__global__ __launch_bounds__(256, 1) void k_batch_gram_oz(const BInst *__restrict__ insts,
                                                          const GramItem *__restrict__ items, int nitems)
{
    __shared__ __attribute__((aligned(16))) unsigned char lds[NBUF][oz::Oz3<S>::STAGE];
    oz::gram_tiles32<S, D, NBUF, LONG>(OzItemSeq{insts, items, nitems}, lds);
}

// Reduce items of instances whose occupied band count is NV (0: any count);
// NEG: instances with a negative frequency (generic count).  One block per
// item (the class's items in decreasing cost, which the dispatcher hands out
// in order): without a persistent item loop the kernels fit three waves per
// SIMD, as the single-problem reduce does (a persistent loop took 169-186
// VGPRs, two waves); the negative-frequency class keeps two above four
// frequencies (it spilled 75 registers at three), and passes of 16 (asked
// for with GPP_OPT_REDUCE_FREQS) run two, as in the single-problem reduce.
template <int FC, int NV, bool NEG = false>
__global__ __launch_bounds__(kRedThreads, FC > 8 || (NEG && FC > 4) ? 2 : 3) void k_batch_reduce(const BInst *__restrict__ insts,
                                                                 const RedItem *__restrict__ items, int chunk)
{
    __shared__ cplx scratch[kRedThreads / kWave];
    const RedItem it = items[blockIdx.x];
    const BInst &b = insts[it.inst];
    ReduceArgs ra;
    ra.mu = b.mu; ra.ieps = b.ieps; ra.wtilde = b.wtilde; ra.aqsn = b.aqsn; ra.bg = b.bg; ra.idx = b.idx;
    ra.vcoul = b.vcoul; ra.wx = b.wx; ra.nv = b.nv; ra.ncouls = b.ncouls; ra.ngpown = b.ngpown;
    ra.nfreq = b.nfreq; ra.nsplit = b.nsplit; ra.chunk = chunk; ra.occ = b.occ;
    ra.col_ach = b.col_ach; ra.col_asx = b.col_asx;
    ra.p2 = nullptr; ra.br1 = nullptr; ra.acht_iw = -1; ra.part_occ = nullptr;
    ra.fpb = b.nfreq;
    pole_reduce_item<FC, NV, false, (FC > 4 ? 4 : FC), NEG>(ra, it.j, it.split, 0, b.nfreq, scratch);
}

__global__ __launch_bounds__(kRedThreads) void k_batch_final(const BInst *__restrict__ insts)
{
    __shared__ cplx scratch[kRedThreads / kWave];
    const BInst &b = insts[blockIdx.x];
    if ((int)blockIdx.y > b.nfreq) return;
    final_block(b.col_ach, b.col_asx, b.part_achs, b.ngpown * b.nsplit, b.nfreq, b.nb, b.nv, b.out, scratch,
                blockIdx.y);
}

__host__ inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

constexpr int gram_tj(int tile) { return tile == 0 || tile == 2 ? 128 : 64; }
constexpr int gram_ti(int tile) { return tile == 0 || tile == 1 ? 128 : 64; }

}  // namespace

// ---------------------------------------------------------------------------
// Reduce item classes: class c in 1..8 holds the instances with c occupied
// bands (c >= 2 run kernels specialised on that count), class 0 the rest,
// class 9 the instances with a negative frequency (any count).
constexpr int kRedClasses = 10;
constexpr int kRedClassNeg = 9;
// Frequencies per reduce pass of a batch (GPP_OPT_REDUCE_FREQS: 4, 16, else 8).
inline int batch_fc(int red_fc) { return red_fc == 4 ? 4 : red_fc == 16 ? 16 : 8; }
// Frequency slots one reduce item evaluates in pole_reduce_item's passes of
// fc, idle slots of the last passes included.
inline int reduce_slots(int nfreq, int fc)
{
    if (fc == 4 || nfreq < 4) return (nfreq + 3) / 4 * 4;
    int slots = 0, r = nfreq;
    if (fc == 16) {
        for (; r > 12; r -= 16) slots += 16;
        if (r > 4) slots += 8, r -= 8;
        return r > 0 ? slots + 4 : slots;
    }
    for (; r > 8; r -= 8) slots += 8;
    return slots + (r > 6 ? 8 : r > 4 ? 6 : r > 2 ? 4 : r > 0 ? 2 : 0);
}
inline int red_class(const DevProblem &p) { return p.wx_neg ? kRedClassNeg : p.nv >= 2 && p.nv <= 8 ? p.nv : 0; }

BatchPlan plan_batch(const std::vector<DevProblem> &probs, const GramKind &k, size_t *work_bytes)
{
    BatchPlan pl;
    pl.n = (int)probs.size();
    for (const DevProblem &p : probs) {
        pl.max_nfreq = std::max(pl.max_nfreq, p.nfreq);
        pl.max_ngpown = std::max(pl.max_ngpown, p.ngpown);
    }
    pl.kind = k;
    size_t off = 0;
    pl.inst_off.resize(pl.n);
    for (int i = 0; i < pl.n; ++i) {
        const DevProblem &p = probs[i];
        const int nsplit = (p.ncouls + k.red_chunk - 1) / k.red_chunk;
        BatchPlan::InstOff o;
        o.nsplit = nsplit;
        o.bg = off;        off += align256(sizeof(cplx) * (size_t)p.nb * p.ngpown);
        o.mu = off;        off += align256(sizeof(cplx) * (size_t)p.ngpown * p.ncouls);
        o.col_ach = off;   off += align256(sizeof(cplx) * (size_t)p.ngpown * nsplit * p.nfreq);
        o.col_asx = off;   off += align256(sizeof(cplx) * (size_t)p.ngpown * nsplit * p.nfreq);
        o.part_achs = off; off += align256(sizeof(cplx) * ach_region_entries(p.nb, p.ngpown));
        o.out = 0;  // in the results table below
        o.oz_eb = o.oz_ea = o.oz_bd = o.oz_ad = 0;
        if (k.ozaki_slices) {
            const oz::OzLayout L = oz::oz_layout(p, k.ozaki_slices, k.int8);
            o.oz_eb = off + L.eb;
            o.oz_ea = off + L.ea;
            o.oz_bd = off + L.bd;
            o.oz_ad = off + L.ad;
            off += align256(L.total);
        }
        pl.inst_off[i] = o;
    }
    // every instance's results (achtemp, asxtemp, per-band terms, achstemp)
    // back to back, so one copy brings them all to the host
    pl.tab_out = off;
    for (int i = 0; i < pl.n; ++i) {
        pl.inst_off[i].out = off;
        off += sizeof(cplx) * ((size_t)2 * probs[i].nfreq + probs[i].nb + 1);
    }
    pl.out_bytes = off - pl.tab_out;
    off = align256(off);
    // device tables after the per-instance areas
    pl.tab_inst = off;  off += align256(sizeof(BInst) * pl.n);
    // work lists
    std::vector<std::pair<long, GramItem>> g;
    std::vector<std::pair<long, RedItem>> r[kRedClasses];
    std::vector<int2> a;
    const int tj = gram_tj(k.tile), ti = gram_ti(k.tile);
    for (int i = 0; i < pl.n; ++i) {
        const DevProblem &p = probs[i];
        const long K = p.nb - p.nv;
        for (int i0 = 0; i0 < p.ncouls; i0 += ti)
            for (int j0 = 0; j0 < p.ngpown; j0 += tj) g.push_back({K, GramItem{i, j0, i0, 0}});
        // One FC for the whole launch (at most 8 for the negative-frequency
        // class): an item runs all frequencies of its (column, G range) in
        // passes of FC with narrower last passes (masked slots beyond nfreq
        // still cost).
        const int fc = p.wx_neg ? std::min(batch_fc(k.red_fc), 8) : batch_fc(k.red_fc);
        const int slots = reduce_slots(p.nfreq, fc);
        const int nsplit = pl.inst_off[i].nsplit;
        for (int j = 0; j < p.ngpown; ++j)
            for (int sp = 0; sp < nsplit; ++sp) {
                const long len = std::min(k.red_chunk, p.ncouls - sp * k.red_chunk);
                r[red_class(p)].push_back({len * slots * (p.nv > 0 ? 3 : 1), RedItem{i, j, sp}});
            }
        for (int t = 0; t < p.nb * kAchSplit; ++t) a.push_back(int2{i, t});
    }
    auto by_cost = [](const auto &x, const auto &y) { return x.first > y.first; };
    std::stable_sort(g.begin(), g.end(), by_cost);
    pl.gram_items.clear();
    for (auto &e : g) pl.gram_items.push_back(e.second);
    for (int c = 0; c < kRedClasses; ++c) {
        std::stable_sort(r[c].begin(), r[c].end(), by_cost);
        pl.red_class_start[c] = (int)pl.red_items.size();
        for (auto &e : r[c]) pl.red_items.push_back(e.second);
    }
    pl.red_class_start[kRedClasses] = (int)pl.red_items.size();
    pl.ach_items = a;
    pl.oz_exp_items.clear();
    pl.oz_split_items.clear();
    pl.oz_gram_items.clear();
    pl.oz_long = false;
    if (k.ozaki_slices) {
        const int KP = k.int8 ? 16 : 8;  // bands per 16-byte digit piece
        std::vector<std::pair<long, GramItem>> og;
        for (int i = 0; i < pl.n; ++i) {
            const oz::OzLayout L = oz::oz_layout(probs[i], k.ozaki_slices, k.int8);
            for (int side = 0; side < 2; ++side) {
                const int npad = side ? L.nc_pad : L.ngp_pad;
                for (int cb = 0; cb < npad / 64; ++cb) pl.oz_exp_items.push_back(int4{i, side, cb, 0});
                for (int cb = 0; cb < (npad + 255) / 256; ++cb)
                    for (int kq = 0; kq < L.kpad / KP; ++kq) pl.oz_split_items.push_back(int4{i, side, cb, kq});
            }
            if (k.int8 && L.kpad > oz::OzDigit<int8_t>::CHUNK) pl.oz_long = true;
            for (int i0 = 0; i0 < L.nc_pad; i0 += oz::kOz3Tile)
                for (int j0 = 0; j0 < L.ngp_pad; j0 += oz::kOz3Tile) og.push_back({L.kpad, GramItem{i, j0, i0, 0}});
        }
        std::stable_sort(og.begin(), og.end(), by_cost);
        for (auto &e : og) pl.oz_gram_items.push_back(e.second);
    }
    pl.tab_gram = off; off += align256(sizeof(GramItem) * (pl.gram_items.size() + 1));
    pl.tab_red = off;  off += align256(sizeof(RedItem) * (pl.red_items.size() + 1));
    pl.tab_ach = off;  off += align256(sizeof(int2) * (pl.ach_items.size() + 1));
    pl.tab_oz_exp = off;   off += align256(sizeof(int4) * (pl.oz_exp_items.size() + 1));
    pl.tab_oz_split = off; off += align256(sizeof(int4) * (pl.oz_split_items.size() + 1));
    pl.tab_oz_gram = off;  off += align256(sizeof(GramItem) * (pl.oz_gram_items.size() + 1));
    pl.total = off;
    pl.fc = batch_fc(k.red_fc);
    if (work_bytes) *work_bytes = pl.total;
    return pl;
}

hipError_t upload_batch(const BatchPlan &pl, const std::vector<DevProblem> &probs, void *work, hipStream_t s)
{
    char *w = static_cast<char *>(work);
    std::vector<BInst> tab(pl.n);
    for (int i = 0; i < pl.n; ++i) {
        const DevProblem &p = probs[i];
        const BatchPlan::InstOff &o = pl.inst_off[i];
        BInst &b = tab[i];
        b.aqsm = p.aqsm; b.aqsn = p.aqsn; b.ieps = p.ieps; b.wtilde = p.wtilde; b.idx = p.idx;
        b.vcoul = p.vcoul; b.wx = p.wx;
        b.nb = p.nb; b.nv = p.nv; b.ncouls = p.ncouls; b.ngpown = p.ngpown; b.nfreq = p.nfreq;
        b.nsplit = o.nsplit; b.occ = p.occ;
        b.bg = reinterpret_cast<cplx *>(w + o.bg);
        b.mu = reinterpret_cast<cplx *>(w + o.mu);
        b.col_ach = reinterpret_cast<cplx *>(w + o.col_ach);
        b.col_asx = reinterpret_cast<cplx *>(w + o.col_asx);
        b.part_achs = reinterpret_cast<cplx *>(w + o.part_achs);
        b.out = reinterpret_cast<cplx *>(w + o.out);
        b.oz_bd = b.oz_ad = nullptr;
        b.oz_eb = b.oz_ea = nullptr;
        b.oz_ngp_pad = b.oz_nc_pad = b.oz_kpad = 0;
        if (pl.kind.ozaki_slices) {
            const oz::OzLayout L = oz::oz_layout(p, pl.kind.ozaki_slices, pl.kind.int8);
            b.oz_bd = w + o.oz_bd;
            b.oz_ad = w + o.oz_ad;
            b.oz_eb = reinterpret_cast<int *>(w + o.oz_eb);
            b.oz_ea = reinterpret_cast<int *>(w + o.oz_ea);
            b.oz_ngp_pad = L.ngp_pad;
            b.oz_nc_pad = L.nc_pad;
            b.oz_kpad = L.kpad;
        }
    }
    hipError_t e;
    if ((e = hipMemcpyAsync(w + pl.tab_inst, tab.data(), sizeof(BInst) * pl.n, hipMemcpyHostToDevice, s))) return e;
    if (!pl.gram_items.empty() &&
        (e = hipMemcpyAsync(w + pl.tab_gram, pl.gram_items.data(), sizeof(GramItem) * pl.gram_items.size(),
                            hipMemcpyHostToDevice, s)))
        return e;
    if (!pl.red_items.empty() &&
        (e = hipMemcpyAsync(w + pl.tab_red, pl.red_items.data(), sizeof(RedItem) * pl.red_items.size(),
                            hipMemcpyHostToDevice, s)))
        return e;
    if (!pl.ach_items.empty() &&
        (e = hipMemcpyAsync(w + pl.tab_ach, pl.ach_items.data(), sizeof(int2) * pl.ach_items.size(),
                            hipMemcpyHostToDevice, s)))
        return e;
    if (!pl.oz_exp_items.empty() &&
        ((e = hipMemcpyAsync(w + pl.tab_oz_exp, pl.oz_exp_items.data(), sizeof(int4) * pl.oz_exp_items.size(),
                             hipMemcpyHostToDevice, s)) ||
         (e = hipMemcpyAsync(w + pl.tab_oz_split, pl.oz_split_items.data(),
                             sizeof(int4) * pl.oz_split_items.size(), hipMemcpyHostToDevice, s)) ||
         (e = hipMemcpyAsync(w + pl.tab_oz_gram, pl.oz_gram_items.data(),
                             sizeof(GramItem) * pl.oz_gram_items.size(), hipMemcpyHostToDevice, s))))
        return e;
    return hipStreamSynchronize(s);
}

template <int S, typename D>
void launch_batch_oz_s(int G, hipStream_t s, const BInst *insts, const int4 *si, int ns, const GramItem *gi, int ng,
                       bool long_range)
{
    constexpr int NBUF = oz::oz3_buffers(S);
    // the split streams memory: up to eight blocks per CU in flight (G is two)
    hipLaunchKernelGGL((k_batch_oz_split<S, D>), dim3(std::max(1, std::min(4 * G, ns))), dim3(256), 0, s, insts, si,
                       ns);
    // one 256-thread block per CU: G is two blocks per CU
    const dim3 grid(std::max(1, std::min((G + 1) / 2, ng)));
    if (long_range)
        hipLaunchKernelGGL((k_batch_gram_oz<S, D, NBUF, true>), grid, dim3(256), 0, s, insts, gi, ng);
    else
        hipLaunchKernelGGL((k_batch_gram_oz<S, D, NBUF, false>), grid, dim3(256), 0, s, insts, gi, ng);
}

// long_range: some instance's band range exceeds one exact INT8 chunk, so the
// tiles flush per chunk (identical results for the shorter ranges).
template <typename D>
void launch_batch_oz(int S, int G, hipStream_t s, const BInst *insts, const int4 *si, int ns, const GramItem *gi,
                     int ng, bool long_range)
{
    switch (S) {
    case 3: launch_batch_oz_s<3, D>(G, s, insts, si, ns, gi, ng, long_range); break;
    case 4: launch_batch_oz_s<4, D>(G, s, insts, si, ns, gi, ng, long_range); break;
    case 5: launch_batch_oz_s<5, D>(G, s, insts, si, ns, gi, ng, long_range); break;
    default: launch_batch_oz_s<6, D>(G, s, insts, si, ns, gi, ng, long_range); break;
    }
}

template <int FC>
void launch_batch_reduce(int cls, dim3 grid, hipStream_t s, const BInst *insts, const RedItem *ri, int chunk)
{
    switch (cls) {
    case 2: hipLaunchKernelGGL((k_batch_reduce<FC, 2>), grid, dim3(kRedThreads), 0, s, insts, ri, chunk); break;
    case 3: hipLaunchKernelGGL((k_batch_reduce<FC, 3>), grid, dim3(kRedThreads), 0, s, insts, ri, chunk); break;
    case 4: hipLaunchKernelGGL((k_batch_reduce<FC, 4>), grid, dim3(kRedThreads), 0, s, insts, ri, chunk); break;
    case 5: hipLaunchKernelGGL((k_batch_reduce<FC, 5>), grid, dim3(kRedThreads), 0, s, insts, ri, chunk); break;
    case 6: hipLaunchKernelGGL((k_batch_reduce<FC, 6>), grid, dim3(kRedThreads), 0, s, insts, ri, chunk); break;
    case 7: hipLaunchKernelGGL((k_batch_reduce<FC, 7>), grid, dim3(kRedThreads), 0, s, insts, ri, chunk); break;
    case 8: hipLaunchKernelGGL((k_batch_reduce<FC, 8>), grid, dim3(kRedThreads), 0, s, insts, ri, chunk); break;
    case kRedClassNeg:  // at most 8 per pass (the 16-wide variant spills)
        hipLaunchKernelGGL((k_batch_reduce<(FC > 8 ? 8 : FC), 0, true>), grid, dim3(kRedThreads), 0, s, insts, ri, chunk);
        break;
    default: hipLaunchKernelGGL((k_batch_reduce<FC, 0>), grid, dim3(kRedThreads), 0, s, insts, ri, chunk); break;
    }
}

// Launches the batch's stages; `which` selects stages (bit 0 gather+achstemp,
// bit 1 Gram, bit 2 reduce, bit 3 final) so the autotuner can time one stage.
hipError_t launch_batch(const BatchPlan &pl, void *work, int persistent_blocks, hipStream_t s, int which)
{
    char *w = static_cast<char *>(work);
    const BInst *insts = reinterpret_cast<const BInst *>(w + pl.tab_inst);
    const int G = persistent_blocks;
    if (which & 1) {
        if (pl.max_ngpown > 0)
            hipLaunchKernelGGL(k_batch_ach_cols, dim3((pl.max_ngpown + 255) / 256, pl.n), dim3(256), 0, s, insts);
        hipLaunchKernelGGL(k_batch_achstemp, dim3(std::min<size_t>(4 * G, pl.ach_items.size() + 1)), dim3(kRedThreads), 0,
                           s, insts, reinterpret_cast<const int2 *>(w + pl.tab_ach), (int)pl.ach_items.size());
    }
    if ((which & 2) && pl.kind.ozaki_slices) {
        const int4 *ei = reinterpret_cast<const int4 *>(w + pl.tab_oz_exp);
        const int4 *si = reinterpret_cast<const int4 *>(w + pl.tab_oz_split);
        const GramItem *gi = reinterpret_cast<const GramItem *>(w + pl.tab_oz_gram);
        const int ne = (int)pl.oz_exp_items.size(), ns = (int)pl.oz_split_items.size();
        const int ng = (int)pl.oz_gram_items.size();
        hipLaunchKernelGGL(k_batch_oz_colexp, dim3(std::max(1, std::min(4 * G, ne))), dim3(256), 0, s, insts, ei,
                           ne);
        if (pl.kind.int8)
            launch_batch_oz<int8_t>(pl.kind.ozaki_slices, G, s, insts, si, ns, gi, ng, pl.oz_long);
        else
            launch_batch_oz<_Float16>(pl.kind.ozaki_slices, G, s, insts, si, ns, gi, ng, false);
    } else if (which & 2) {
        const GramItem *gi = reinterpret_cast<const GramItem *>(w + pl.tab_gram);
        const int ng = (int)pl.gram_items.size();
        const dim3 grid(std::max(1, std::min(G, ng)));
        switch (pl.kind.tile) {
        case 0: hipLaunchKernelGGL((k_batch_gram<2, 4>), grid, dim3(512), 0, s, insts, gi, ng); break;
        case 1: hipLaunchKernelGGL((k_batch_gram<1, 4>), grid, dim3(256), 0, s, insts, gi, ng); break;
        case 2: hipLaunchKernelGGL((k_batch_gram<2, 2>), grid, dim3(256), 0, s, insts, gi, ng); break;
        default: hipLaunchKernelGGL((k_batch_gram<1, 2>), grid, dim3(128), 0, s, insts, gi, ng); break;
        }
    }
    if (which & 4) {
        // One launch per occupied band class present, one block per item.
        for (int c = 0; c < kRedClasses; ++c) {
            const int b0 = pl.red_class_start[c], nr = pl.red_class_start[c + 1] - b0;
            if (nr == 0) continue;
            const RedItem *ri = reinterpret_cast<const RedItem *>(w + pl.tab_red) + b0;
            const dim3 grid(nr);  // one block per item
            if (pl.fc == 4)
                launch_batch_reduce<4>(c, grid, s, insts, ri, pl.kind.red_chunk);
            else if (pl.fc == 16)
                launch_batch_reduce<16>(c, grid, s, insts, ri, pl.kind.red_chunk);
            else
                launch_batch_reduce<8>(c, grid, s, insts, ri, pl.kind.red_chunk);
        }
    }
    if (which & 8)
        hipLaunchKernelGGL(k_batch_final, dim3(pl.n, pl.max_nfreq + 1), dim3(kRedThreads), 0, s, insts);
    return hipGetLastError();
}

size_t batch_results_offset(const BatchPlan &pl, int i) { return pl.inst_off[i].out; }

}  // namespace gpp
