// gpp_device.hpp -- device-side complex type, reference-order complex helpers
// and deterministic wavefront / block reductions shared by the GPP kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gpp {

constexpr int kWave = 64;  // CDNA wavefront width

// 16-byte interleaved complex, layout-identical to the reference's GPPComplex
// (Complex.h:15-19) and to gpp_complex of include/gpp.h.
struct alignas(16) cplx {
    double re;
    double im;
};

__host__ __device__ __forceinline__ cplx mk(double re, double im) { return cplx{re, im}; }
__host__ __device__ __forceinline__ cplx czero() { return cplx{0.0, 0.0}; }

// The helpers below follow the reference's operation order and are compiled
// without FMA contraction, so each rounds exactly as the x86-64 reference
// does.  The pragma sits in every body, the multiplications and additions
// alone included: it does not reach inlined callees, and an unmarked multiply
// of one helper feeding an unmarked add of another (delw = x * rden, then
// 0.5 + delw) is otherwise fused into one FMA.

// GPPComplex_product, Complex.h:167-174.
__device__ __forceinline__ cplx cmul(cplx a, cplx b)
{
#pragma clang fp contract(off)
    return cplx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// GPPComplex_conj, Complex.h:153-161.
__device__ __forceinline__ cplx cconj(cplx a)
{
#pragma clang fp contract(off)
    return cplx{a.re, -1 * a.im};
}
// GPPComplex_square, Complex.h:141-148.
__device__ __forceinline__ cplx csquare(cplx a)
{
#pragma clang fp contract(off)
    return cplx{a.re * a.re - a.im * a.im, 2 * a.re * a.im};
}
// GPPComplex_mult(a, b), Complex.h:200-205.
__device__ __forceinline__ cplx cscale(cplx a, double b)
{
#pragma clang fp contract(off)
    return cplx{a.re * b, a.im * b};
}
// GPPComplex_mult(a, b, c), Complex.h:190-195: ((re*b)*c, (im*b)*c).
__device__ __forceinline__ cplx cscale2(cplx a, double b, double c)
{
#pragma clang fp contract(off)
    return cplx{a.re * b * c, a.im * b * c};
}
// doubleMinusGPPComplex, Complex.h:234-237.
__device__ __forceinline__ cplx cdminus(double a, cplx z)
{
#pragma clang fp contract(off)
    return cplx{a - z.re, 0 - z.im};
}
// doublePlusGPPComplex, Complex.h:239-242.
__device__ __forceinline__ cplx cdplus(double a, cplx z)
{
#pragma clang fp contract(off)
    return cplx{a + z.re, 0 + z.im};
}
// GPPComplex_abs, Complex.h:179-185 (with sqrt).
__device__ __forceinline__ double cabs_(cplx a)
{
#pragma clang fp contract(off)
    return sqrt(a.re * a.re + a.im * a.im);
}
// operator+=, Complex.h:51-56: dst = src + dst.
__device__ __forceinline__ void cacc(cplx &dst, cplx src)
{
#pragma clang fp contract(off)
    dst.re = src.re + dst.re;
    dst.im = src.im + dst.im;
}
__device__ __forceinline__ cplx cadd(cplx a, cplx b)
{
#pragma clang fp contract(off)
    return cplx{a.re + b.re, a.im + b.im};
}

// 1/d for d > 0 to within about one ulp: the hardware estimate refined by two
// Newton-Raphson steps (each at least doubles the correct bits).  d = 0 gives
// NaN, as the reference's 1/0 = inf times the zero numerator it multiplies;
// NaN propagates.  Valid for d in [2^-1000, 2^1000]; beyond that range the
// result is NaN (callers needing IEEE behaviour there use the strict form).
__device__ __forceinline__ double rcp_refined(double d)
{
    double r = __builtin_amdgcn_rcp(d);
    double e = __builtin_fma(-d, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-d, r, 1.0);
    r = __builtin_fma(r, e, r);
    return r;
}

// 1/d with one Newton-Raphson step: about 2^-45 relative (the estimate is
// good to about 2^-23).  For values that feed sums, not branch decisions.
__device__ __forceinline__ double rcp_1nr(double d)
{
    double r = __builtin_amdgcn_rcp(d);
    const double e = __builtin_fma(-d, r, 1.0);
    return __builtin_fma(r, e, r);
}

// 1/d as r0 (1 + e + e^2) with r0 the hardware estimate and e = 1 - d r0:
// the relative error is e^3 (about 2^-70) plus the final rounding, so it is
// unbiased like a correctly rounded quotient (one Newton step leaves -e^2,
// always of one sign, which a long sum with cancellation accumulates).  One
// FMA more than rcp_1nr.  d = 0 gives NaN (0 * inf in the correction).
__device__ __forceinline__ double rcp_2o(double d)
{
    const double r = __builtin_amdgcn_rcp(d);
    const double e = __builtin_fma(-d, r, 1.0);
    return __builtin_fma(r, __builtin_fma(e, e, e), r);
}

// A wave-uniform value moved into scalar registers (readfirstlane of each
// 32-bit half): it then costs no VGPRs where it is live.  Only for values
// every lane of the wave holds alike.
__device__ __forceinline__ double uniform(double x)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ cplx uniform(cplx x) { return cplx{uniform(x.re), uniform(x.im)}; }

// Product a*b with FMA contraction allowed (fast paths only).
__device__ __forceinline__ cplx cmul_fast(cplx a, cplx b)
{
    return cplx{__builtin_fma(a.re, b.re, -a.im * b.im), __builtin_fma(a.re, b.im, a.im * b.re)};
}

// acc + a*b with four FMAs.
__device__ __forceinline__ cplx cfma(cplx a, cplx b, cplx acc)
{
    return cplx{__builtin_fma(a.re, b.re, __builtin_fma(-a.im, b.im, acc.re)),
                __builtin_fma(a.re, b.im, __builtin_fma(a.im, b.re, acc.im))};
}

// XCD-aware tile order.  The dispatcher hands workgroup b (linear id) to XCD
// b % 8, each XCD with its own L2; tile t of nt is given to the workgroup such
// that every XCD walks a contiguous range of tiles, so neighbouring tiles
// (which share an operand tile) meet in one L2.  A bijection on [0, nt).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_tile(int b, int nt)
{
    const int q = nt / kXcds, r = nt % kXcds;
    const int x = b % kXcds, i = b / kXcds;
    return x * q + (x < r ? x : r) + i;
}

// Wave-wide sum by xor butterfly.  Every lane ends with the same, bitwise
// reproducible value (a+b == b+a in IEEE arithmetic).
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int m = 1; m < kWave; m <<= 1) v += __shfl_xor(v, m, kWave);
    return v;
}
__device__ __forceinline__ cplx wave_sum(cplx v) { return cplx{wave_sum(v.re), wave_sum(v.im)}; }

// Block-wide sum in a fixed order: butterfly within each wave, then the wave
// totals in wave order.  Every thread of the block returns the same value.
// `scratch` needs blockDim.x / 64 entries; the call contains two barriers.
__device__ __forceinline__ cplx block_sum(cplx v, cplx *scratch)
{
    const int nw = blockDim.x / kWave;
    const int w = threadIdx.x / kWave;
    v = wave_sum(v);
    if ((threadIdx.x % kWave) == 0) scratch[w] = v;
    __syncthreads();
    cplx r = czero();
    for (int i = 0; i < nw; ++i) r = cadd(r, scratch[i]);
    __syncthreads();
    return r;
}

constexpr double kTo1 = 1e-6;    // to1, gppComplex.cpp:20, :241
constexpr double kSexcut = 4.0;  // sexcut, gppComplex.cpp:21, :243

// Occupied-band pole terms before the ssx cutoff, ssxt_scht_solver
// (gppComplex.cpp:16-56), in the reference's operation order with IEEE
// division and no contraction, so the branch decisions and values are the
// reference's.  wx2 = wxt^2 and rwx4 = 1/wx2^2 (:38-40) are passed in.
__device__ __forceinline__ void occ_pole_terms_uncut(double wxt, double wx2, double rwx4, cplx wt, cplx ie,
                                                     cplx &sch, cplx &ssx)
{
#pragma clang fp contract(off)
    const double limitone = 1.0 / (kTo1 * 4.0);  // :22
    const double limittwo = 0.25;                // :23, pow(0.5,2)
    const cplx wt2 = csquare(wt);                // flagOCC_solver :117
    const cplx om2 = cmul(wt2, ie);              // :118
    sch = czero();
    ssx = czero();
    const cplx wdiff = cdminus(wxt, wt);                         // :27
    const double rden = 1 / cmul(wdiff, cconj(wdiff)).re;        // :30
    const cplx delw = cscale(cmul(wt, cconj(wdiff)), rden);      // :31
    const double delwr = cmul(delw, cconj(delw)).re;             // :32
    const double wdiffr = cmul(wdiff, cconj(wdiff)).re;          // :33
    if ((wdiffr > limittwo) && (delwr < limitone)) {             // :35
        sch = cmul(delw, ie);                                    // :37
        ssx = cscale2(om2, wx2, rwx4);                           // :38-41
    } else if (delwr > kTo1) {                                   // :43
        const cplx cden = cscale(cmul(wt2, cdplus(0.50, delw)), 4.00);  // :46
        const double rd = 1.00 / cmul(cden, cconj(cden)).re;            // :48-49
        ssx = cmul(cmul(cplx{-om2.re, -om2.im}, cconj(cden)), cscale(delw, rd));  // :50
    }
}

// |ssx| of the reference's branch 2 (gppComplex.cpp:44-50) for one element,
// in its operation order: the rare near-cutoff path of the factored reduce
// (only in its negative-frequency kernels).
__device__ __forceinline__ double occ_branch2_abs_ssx(double wxt, cplx wt, cplx ie)
{
#pragma clang fp contract(off)
    const cplx wt2 = csquare(wt);                                          // :117
    const cplx om2 = cmul(wt2, ie);                                        // :118
    const cplx wdiff = cdminus(wxt, wt);                                   // :27
    const double rden = 1 / cmul(wdiff, cconj(wdiff)).re;                  // :30
    const cplx delw = cscale(cmul(wt, cconj(wdiff)), rden);                // :31
    const cplx cden = cscale(cmul(wt2, cdplus(0.50, delw)), 4.00);         // :46
    const double rd = 1.00 / cmul(cden, cconj(cden)).re;                   // :48-49
    return cabs_(cmul(cmul(cplx{-om2.re, -om2.im}, cconj(cden)), cscale(delw, rd)));  // :50, :59
}

// ssxt_scht_solver (gppComplex.cpp:16-63) with the ssx cutoff of :58-59
// (only reachable for wx < 0).
__device__ __forceinline__ void occ_pole_terms(double wxt, double wx2, double rwx4, cplx wt, cplx ie,
                                               cplx &sch, cplx &ssx)
{
#pragma clang fp contract(off)
    occ_pole_terms_uncut(wxt, wx2, rwx4, wt, ie, sch, ssx);
    if (wxt < 0.00) {
        const double ssxcutoff = cabs_(ie) * kSexcut;
        if (cabs_(ssx) > ssxcutoff) ssx = czero();
    }
}

// reduce_achstemp (gppComplex.cpp:66-101) for band n1 by one wave: lane l
// sums columns l, l + 64, ... in order, then the xor butterfly (every lane
// gets the same, bitwise reproducible total).  With bg non-null it also
// stores the gathered row Bg[n1][j] = aqsm[n1][igp(j)] the factored forms
// use: the scattered column reads of aqsm are shared by both.
// The else branch of :94-98 is unreachable once negative indices are refused.
// Each band's achstemp sum is split over this many waves (columns j with
// j / 64 = q mod kAchSplit go to wave q); the partials are kept per (band, q)
// and summed in the final stage.  One wave per band left most SIMDs idle
// and each wave waited on a chain of dependent scattered loads.
constexpr int kAchSplit = 4;

// The band-independent part of column j's achstemp term (gppComplex.cpp:
// 77-99): igp, schs = -I_eps[j][igp] when |schs| > to1 (else zero: the term
// is skipped) and vcoul[igp] * 0.5, formed once per call instead of once per
// band (the strided I_eps element and the square root were the costly part).
struct AchCol {
    double sre, sim, vh;
    int igp, pad;
};
// The achstemp region of a workspace, in complex entries: kAchSplit partials
// per band, then (from an even entry, 32-byte aligned) one AchCol per column.
__host__ __device__ inline size_t ach_cols_entry(int nb) { return ((size_t)nb * kAchSplit + 2) & ~(size_t)1; }
__host__ __device__ inline size_t ach_region_entries(int nb, int ngpown)
{
    return ach_cols_entry(nb) + 2 * (size_t)ngpown;
}

__device__ __forceinline__ void ach_col(const cplx *ieps, const int32_t *idx, const double *vcoul, int ncouls, int j,
                                        AchCol *cols);
__device__ __forceinline__ cplx achstemp_band(const cplx *aqsm, const cplx *aqsn, const AchCol *cols, int ncouls,
                                               int ngpown, int n1, int q, cplx *bg);

// igp = min(inv_igp_index[j], ncouls-1), gppComplex.cpp:342-344.  Negative
// entries are rejected before any launch; the max(.,0) only keeps a kernel
// in bounds should one slip through.
__device__ __forceinline__ int igp_of(const int32_t *idx, int j, int ncouls)
{
    int igp = idx[j];
    igp = igp >= ncouls ? ncouls - 1 : igp;
    return igp < 0 ? 0 : igp;
}

__device__ __forceinline__ void ach_col(const cplx *ieps, const int32_t *idx, const double *vcoul, int ncouls, int j,
                                        AchCol *cols)
{
    const int igp = igp_of(idx, j, ncouls);
    const cplx ie = ieps[(size_t)j * ncouls + igp];
    const cplx schs = cplx{-ie.re, -ie.im};  // :86-88
    const bool keep = cabs_(schs) > kTo1;    // :91
    cols[j] = AchCol{keep ? schs.re : 0.0, keep ? schs.im : 0.0, vcoul[igp] * 0.5, igp, 0};
}

__device__ __forceinline__ cplx achstemp_band(const cplx *aqsm, const cplx *aqsn, const AchCol *cols, int ncouls,
                                               int ngpown, int n1, int q, cplx *bg)
{
    cplx s = czero();
    // unrolled: four columns' dependent loads (the column entry, then the
    // scattered aqsm / aqsn elements) in flight per lane
#pragma unroll 4
    for (int j = q * kWave + threadIdx.x % kWave; j < ngpown; j += kAchSplit * kWave) {
        const AchCol c = cols[j];
        const cplx m = aqsm[(size_t)n1 * ncouls + c.igp];
        if (bg) bg[(size_t)n1 * ngpown + j] = m;
        if (c.sre != 0.0 || c.sim != 0.0) {  // |schs| > to1: schs is nonzero
            const cplx matngmatmgp = cmul(aqsn[(size_t)n1 * ncouls + c.igp], cconj(m));  // :84, :89
            cplx schstemp = czero();
            cacc(schstemp, cmul(matngmatmgp, cplx{c.sre, c.sim}));  // :92
            cacc(s, cscale(schstemp, c.vh));                         // :99
        }
    }
    return cplx{wave_sum(s.re), wave_sum(s.im)};
}

}  // namespace gpp
