// gpp_internal.hpp -- host-side interfaces shared between the C ABI layer
// (gpp_api.cpp) and the kernel launchers (gpp_direct.hip, gpp_factored.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "gpp_device.hpp"

namespace gpp {

// One problem's sizes and device inputs (row-major, see include/gpp.h).
struct DevProblem {
    int nb = 0, nv = 0, ncouls = 0, ngpown = 0, nfreq = 0;
    double occ = 1.0;
    const int32_t *idx = nullptr;  // [ngpown]
    const double *wx = nullptr;    // [nfreq], or [nb][nfreq] when wx_stride == nfreq
    int wx_stride = 0;             // 0: one wx row for every band; nfreq: a row per band
    int wx_neg = 0;                // the shared row has a negative frequency (ssx cutoff of gppComplex.cpp:58-59
                                   // live), known on the host; the factored reduce selects its kernel by it
    const double *vcoul = nullptr; // [ncouls]
    const cplx *aqsm = nullptr;    // [nb][ncouls]
    const cplx *aqsn = nullptr;    // [nb][ncouls]
    const cplx *ieps = nullptr;    // [ngpown][ncouls]
    const cplx *wtilde = nullptr;  // [ngpown][ncouls]
};

// Device results of one evaluation, contiguous:
//   achtemp[nfreq], asxtemp[nfreq], acht_n1_loc[nb], achstemp[1].
struct DevResults {
    cplx *base = nullptr;
    int nfreq = 0, nb = 0;
    cplx *achtemp() const { return base; }
    cplx *asxtemp() const { return base + nfreq; }
    cplx *acht_n1_loc() const { return base + 2 * nfreq; }
    cplx *achstemp() const { return base + 2 * nfreq + nb; }
    size_t count() const { return (size_t)2 * nfreq + nb + 1; }
};

// Grow-only device scratch owned by a context (or by the reference-
// compatible launchers' per-device cache).
struct Workspace {
    void *ptr = nullptr;
    size_t bytes = 0;
    int device = -1;
    hipError_t reserve(size_t need);
    void release();
};

// Records HIP events between the stages of an evaluation, for the last
// kRing evaluations (so that evaluations enqueued back to back can be timed
// after one synchronisation).
struct StageTimer {
    static constexpr int kMax = 12;
    static constexpr int kRing = 32;
    struct Slot {
        hipEvent_t ev[kMax + 1] = {};
        const char *name[kMax] = {};
        int n = 0;  // stages recorded
    };
    Slot slot[kRing];
    int cur = 0;     // slot of the current (last) evaluation
    int filled = 0;  // evaluations recorded, at most kRing
    bool ready = false;
    hipError_t init();
    void destroy();
    void begin(hipStream_t s);                 // starts a slot, records ev[0]
    void mark(hipStream_t s, const char *nm);  // closes one stage
    const Slot &last() const { return slot[cur]; }
    const Slot &back(int k) const { return slot[(cur - k + kRing) % kRing]; }  // k evaluations before the last
};

// 1 if some of the n frequencies is negative (NaN counts as negative, so an
// unordered value takes the kernels that check every frequency).
inline int wx_has_negative(const double *wx, int n)
{
    for (int i = 0; i < n; ++i)
        if (!(wx[i] >= 0.0)) return 1;
    return 0;
}

// Bytes of device workspace the direct form needs for this problem.
size_t direct_workspace_bytes(const DevProblem &p);

// Direct FP64 evaluation (strict = reference operation order).  Writes the
// results to `res` (overwritten).  Workspace must hold
// direct_workspace_bytes(p).
hipError_t run_direct(const DevProblem &p, bool strict, void *work, const DevResults &res,
                      hipStream_t s, StageTimer *timer);

// Only the occupied bands (n1 < nv) / only the unoccupied bands of the direct
// form, plus achstemp, for the reference-compatible launchers.  `which`
// selects 0 = occupied, 1 = unoccupied, 2 = achstemp.
hipError_t run_direct_part(const DevProblem &p, int which, void *work, const DevResults &res,
                           hipStream_t s);

// reduce_achstemp per band into part_achs[nb] (fixed-order sums over j).
// achstemp per band into part_achs[nb]; with bg non-null also the gathered
// rows Bg[n1][j] = aqsm[n1][igp(j)] [nb][ngpown].
hipError_t launch_achstemp(const DevProblem &p, cplx *part_achs, cplx *bg, hipStream_t s);

// Zeroes n ints with a fill kernel (no memset node when a caller captures).
hipError_t launch_zero_ints(int *d_p, int n, hipStream_t s);

// Counts negative inv_igp_index entries on the device (result in *d_count).
hipError_t launch_check_index(const int32_t *d_idx, int ngpown, int *d_count, hipStream_t s);

// Adds res (nfreq=3 layout) into reference-style split outputs.
hipError_t launch_accumulate_compat(const DevResults &res, int which, int nv, double *achtemp_re,
                                    double *achtemp_im, cplx *asxtemp, cplx *acht_n1_loc,
                                    double *achstemp_re, double *achstemp_im, hipStream_t s);

// Seeded input generation on the device (counter-based hash, identical bits
// to the host generator).
hipError_t launch_fill_seeded(cplx *dst, size_t n, uint64_t seed, uint32_t array_id, double re_lo,
                              double re_hi, double im_lo, double im_hi, hipStream_t s);

// How the factored form computes its Gram product and reduce.
struct GramKind {
    int ozaki_slices = 0;  // 0: FP64 MFMA; 3..6: Ozaki split with that many slices
    bool int8 = false;     // Ozaki digits in INT8 (else FP16)
    int layout = 4;        // Ozaki block layout (0, 1: 64x64 staged in registers; 2: 64x32; 3: 64x64 on 32x32 MFMAs;
                           // 4: layout 3 in persistent blocks; 2-4 load straight into LDS)
    int group = 0;         // Ozaki G' tiles per tile group (0: sized for the L2)
    int tile = 0;          // FP64 Gram tile: 0 128x128, 1 64x128, 2 128x64, 3 64x64 (G' x G)
    int red_chunk = 8192;  // G-vectors per pole-reduce block
    int red_fc = 0;        // frequencies per pole-reduce pass when nfreq > 4 (4, 8, 16; 0: 16 from 16 up, else 8;
                           // the batched form: 4, else 8)
    int red_fpb = 0;       // frequencies per pole-reduce block (0: all, passes looped in the block)
    bool acht = false;     // also produce acht_n1_loc (a second product, R * aqsn^T)
    bool fuse = false;     // pole reduce in the FP64 Gram epilogue where it applies (no Mu workspace)
};
constexpr int kGramTiles = 4;

size_t factored_workspace_bytes(const DevProblem &p, GramKind kind);
hipError_t run_factored(const DevProblem &p, GramKind kind, void *work, const DevResults &res,
                        hipStream_t s, StageTimer *timer);

// Band-dependent frequencies (gpp_factored.hip): bands grouped by identical
// wx rows; bucket b holds bands perm[start[b] .. start[b]+size[b]), its
// nocc[b] occupied bands first, and uses row `row[b]` (its first band).
struct BucketPlan {
    std::vector<int32_t> perm;
    std::vector<int> start, size, nocc, row;
    std::vector<int> neg;      // per bucket: its row has a negative frequency
    std::vector<int32_t> tab;  // per bucket: result offset, start, size
    size_t nres = 0;           // complex results of all buckets (DevResults each)
};
BucketPlan plan_buckets(const double *wx_rows, int nb, int nv, int nfreq);
size_t bucketed_workspace_bytes(const DevProblem &p, GramKind kind, const BucketPlan &bp);
// p.wx is [nb][nfreq] (wx_stride == nfreq); the host vectors of bp must stay
// alive until the stream has run the call.
hipError_t run_factored_bucketed(const DevProblem &p, GramKind kind, const BucketPlan &bp, void *work,
                                 const DevResults &res, hipStream_t s, StageTimer *timer);

// Ozaki Gram product (gpp_ozaki.hip): Mu = conj(Bg)^T aqsn over bands
// [nv, nb), written to mout [ngpown][ncouls].
size_t ozaki_workspace_bytes(const DevProblem &p, int slices, bool int8);
hipError_t launch_gram_ozaki(const DevProblem &p, int slices, bool int8, int layout, int group, const cplx *bg,
                             void *work, cplx *mout, hipStream_t s);

// ---- batched factored form (gpp_batch.hip) --------------------------------
struct GramItem {  // one Gram tile of one instance
    int inst, j0, i0, pad;
};
struct RedItem {   // one pole-reduce item of one instance: column j, G range split, every frequency
    int inst, j, split;
};

// Workspace layout and work lists of one batch for one GramKind.
struct BatchPlan {
    struct InstOff {
        size_t bg, mu, col_ach, col_asx, part_achs, out;
        size_t oz_eb, oz_ea, oz_bd, oz_ad;  // Ozaki digit areas (Ozaki kinds only)
        int nsplit;
    };
    int n = 0;
    int fc = 8;
    int max_nfreq = 0;
    int max_ngpown = 0;
    GramKind kind;
    std::vector<InstOff> inst_off;
    std::vector<GramItem> gram_items;
    std::vector<RedItem> red_items;  // grouped by occupied band class (plan_batch), each class by cost
    int red_class_start[11] = {};    // items of class c: [red_class_start[c], red_class_start[c + 1])
    std::vector<int2> ach_items;
    // Ozaki kinds: column exponents (inst, side, 64-column block), digit
    // pieces (inst, side, 256-column block, piece) and 64 x 64 Gram tiles.
    std::vector<int4> oz_exp_items, oz_split_items;
    std::vector<GramItem> oz_gram_items;
    bool oz_long = false;  // an INT8 instance's band range exceeds one exact chunk
    size_t tab_out = 0, out_bytes = 0;  // the results table, [2 nfreq + nb + 1] per instance
    size_t tab_inst = 0, tab_gram = 0, tab_red = 0, tab_ach = 0, total = 0;
    size_t tab_oz_exp = 0, tab_oz_split = 0, tab_oz_gram = 0;
};

BatchPlan plan_batch(const std::vector<DevProblem> &probs, const GramKind &k, size_t *work_bytes);

// Writes the device tables of the plan into `work` (synchronous on s).
hipError_t upload_batch(const BatchPlan &pl, const std::vector<DevProblem> &probs, void *work, hipStream_t s);
// Enqueues the batch's persistent stages; which: bit 0 gather+achstemp, bit 1
// Gram, bit 2 reduce, bit 3 final (15 = all).
hipError_t launch_batch(const BatchPlan &pl, void *work, int persistent_blocks, hipStream_t s, int which);
// Offset in `work` of instance i's results (DevResults layout).
size_t batch_results_offset(const BatchPlan &pl, int i);

}  // namespace gpp
