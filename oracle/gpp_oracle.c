/*
 * gpp_oracle.c -- CPU oracle for the GPP sigma band loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the
 * smoke() entry point and bench.py's cpu_baseline leg may load it.  No
 * product code path calls it, and the HIP library never falls back to it.
 *
 * It restates, in plain C, the arithmetic of the reference mini-app's
 * sequential build (gppComplex.cpp with the GPPComplex class of Complex.h),
 * keeping the reference's operation order so that, compiled with
 * -ffp-contract=off, it reproduces the reference bit for bit.  Every
 * function cites the reference lines it follows.
 *
 * Layout: complex arrays are interleaved (re, im) doubles, row-major:
 *   aqsmtemp, aqsntemp  [number_bands][ncouls]
 *   I_eps_array, wtilde [ngpown][ncouls]
 *   vcoul [ncouls], inv_igp_index [ngpown], wx_array [nfreq]
 *
 * Pinning (what checks each function; the tests hold them):
 *   oc_* complex helpers      -- numpy complex128 on random operands; the
 *                                README golden triples through the loop.
 *   oracle_pole_terms         -- each branch against hand-derived closed
 *                                forms and numpy; the compiled reference
 *                                harness (oracle/_ref) on seeded/near-pole
 *                                fixtures.
 *   oracle_sigma (seq/omp)    -- README golden achtemp triples (512 2
 *                                {800,1600,3200} 20), the closed-form integer
 *                                asxtemp, exact-rational brute force on tiny
 *                                shapes, invariants, and the compiled
 *                                reference harness on seeded fixtures.
 *   oracle_sigma_dd           -- its TwoSum / dd-sum / dd-times-double
 *                                helpers against exact rational sums and
 *                                products (oracle_dd_* hooks); the whole
 *                                loop bitwise equal to the correctly rounded
 *                                exact value on dyadic inputs whose terms are
 *                                exact but whose left-to-right sum loses
 *                                bits (where oracle_sigma is measurably
 *                                off), and within 2e-16 of sum|terms| of the
 *                                exact-rational brute force on seeded,
 *                                near-pole and cancelling shapes.
 *   oracle_sigma_wxb          -- equal rows reproduce oracle_sigma bitwise;
 *                                distinct rows: each band's acht_n1_loc equals,
 *                                bitwise, the one-band run with that band's
 *                                row (oracle and compiled reference harness),
 *                                and achtemp/asxtemp/achstemp equal the sum of
 *                                the one-band runs to rounding.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double re, im; } oc_cplx;

int oracle_sigma_wxb(int nb, int nv, int ncouls, int ngpown, int nfreq, double occ,
                     const int *inv_igp_index, const double *wx_array, int wx_stride, const double *vcoul,
                     const double *aqsmtemp, const double *aqsntemp,
                     const double *I_eps_array, const double *wtilde_array,
                     double *achtemp_out, double *asxtemp_out, double *acht_n1_loc_out,
                     double *achstemp_out, int nthreads);

/* GPPComplex(x, y), Complex.h:27-30 */
static inline oc_cplx oc_make(double re, double im) { oc_cplx r; r.re = re; r.im = im; return r; }

/* GPPComplex_product, Complex.h:167-174 */
static inline oc_cplx oc_product(oc_cplx a, oc_cplx b)
{
    return oc_make(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}

/* GPPComplex_conj, Complex.h:153-161 (im = -1 * im) */
static inline oc_cplx oc_conj(oc_cplx a) { return oc_make(a.re, -1 * a.im); }

/* GPPComplex_square, Complex.h:141-148 */
static inline oc_cplx oc_square(oc_cplx a) { return oc_make(a.re * a.re - a.im * a.im, 2 * a.re * a.im); }

/* GPPComplex_abs, Complex.h:179-185 (with sqrt) */
static inline double oc_abs(oc_cplx a) { return sqrt(a.re * a.re + a.im * a.im); }

/* GPPComplex_mult(a, b), Complex.h:200-205 */
static inline oc_cplx oc_mult(oc_cplx a, double b) { return oc_make(a.re * b, a.im * b); }

/* GPPComplex_mult(a, b, c) = ((re*b)*c, (im*b)*c), Complex.h:190-195 */
static inline oc_cplx oc_mult3(oc_cplx a, double b, double c) { return oc_make(a.re * b * c, a.im * b * c); }

/* doubleMinusGPPComplex, Complex.h:234-237 */
static inline oc_cplx oc_dminus(double a, oc_cplx z) { return oc_make(a - z.re, 0 - z.im); }

/* doublePlusGPPComplex, Complex.h:239-242 */
static inline oc_cplx oc_dplus(double a, oc_cplx z) { return oc_make(a + z.re, 0 + z.im); }

/* operator+=, Complex.h:51-56: this = src + this */
static inline void oc_acc(oc_cplx *dst, oc_cplx src)
{
    dst->re = src.re + dst->re;
    dst->im = src.im + dst->im;
}

/* operator- (negation), Complex.h:65-70 */
static inline oc_cplx oc_neg(oc_cplx a) { return oc_make(-a.re, -a.im); }

static inline oc_cplx oc_load(const double *p, long i) { return oc_make(p[2 * i], p[2 * i + 1]); }

/* Constants of the reference, gppComplex.cpp:20-23 and :239-247. */
#define OC_TO1 1e-6
#define OC_SEXCUT 4.0

/*
 * The per-(iw, ig) pole terms of the occupied-band solver,
 * ssxt_scht_solver, gppComplex.cpp:16-63, before the multiplication by the
 * matrix element (lines 61-62).  Returns the branch taken: 1 (lines 35-42),
 * 2 (lines 43-51) or 3 (lines 52-56).  The ssx cutoff of lines 58-59 is
 * applied.
 */
int oracle_pole_terms(double wxt, const double *wtilde_p, const double *ieps_p,
                      double *sch_out, double *ssx_out)
{
    const double limitone = 1.0 / (OC_TO1 * 4.0);
    const double limittwo = pow(0.5, 2);
    oc_cplx wtilde = oc_load(wtilde_p, 0);
    oc_cplx ieps = oc_load(ieps_p, 0);
    /* flagOCC_solver, gppComplex.cpp:117-118 */
    oc_cplx wtilde2 = oc_square(wtilde);
    oc_cplx Omega2 = oc_product(wtilde2, ieps);

    oc_cplx sch = oc_make(0.0, 0.0), ssx = oc_make(0.0, 0.0);
    oc_cplx wdiff = oc_dminus(wxt, wtilde);                              /* :27 */
    oc_cplx cden = wdiff;                                                /* :29 */
    double rden = 1 / oc_product(cden, oc_conj(cden)).re;                /* :30 */
    oc_cplx delw = oc_mult(oc_product(wtilde, oc_conj(cden)), rden);     /* :31 */
    double delwr = oc_product(delw, oc_conj(delw)).re;                   /* :32 */
    double wdiffr = oc_product(wdiff, oc_conj(wdiff)).re;                /* :33 */
    int branch;
    if ((wdiffr > limittwo) && (delwr < limitone)) {                     /* :35 */
        sch = oc_product(delw, ieps);                                    /* :37 */
        double cdend = pow(wxt, 2);                                      /* :38 */
        rden = pow(cdend, 2);                                            /* :39 */
        rden = 1.00 / rden;                                              /* :40 */
        ssx = oc_mult3(Omega2, cdend, rden);                             /* :41 */
        branch = 1;
    } else if (delwr > OC_TO1) {                                         /* :43 */
        cden = oc_mult(oc_product(wtilde2, oc_dplus(0.50, delw)), 4.00); /* :46 */
        rden = oc_product(cden, oc_conj(cden)).re;                       /* :48 */
        rden = 1.00 / rden;                                              /* :49 */
        ssx = oc_product(oc_product(oc_neg(Omega2), oc_conj(cden)), oc_mult(delw, rden)); /* :50 */
        branch = 2;
    } else {
        branch = 3;                                                      /* :52-56 */
    }
    double ssxcutoff = oc_abs(ieps) * OC_SEXCUT;                         /* :58 */
    if ((oc_abs(ssx) > ssxcutoff) && (wxt < 0.00)) ssx = oc_make(0.0, 0.0); /* :59 */
    sch_out[0] = sch.re; sch_out[1] = sch.im;
    ssx_out[0] = ssx.re; ssx_out[1] = ssx.im;
    return branch;
}

/*
 * flagOCC_solver + ssxt_scht_solver for one (n1, my_igp, iw):
 * gppComplex.cpp:107-128 and :16-63.  Returns ssxt, scht.
 */
static void oc_flag_occ(double wxt, const double *wtilde_array, int my_igp, int n1,
                        const double *aqsmtemp, const double *aqsntemp, const double *I_eps_array,
                        oc_cplx *ssxt, oc_cplx *scht, int ncouls, int igp)
{
    const long row = (long)my_igp * ncouls;
    const long brow = (long)n1 * ncouls;
    for (int ig = 0; ig < ncouls; ++ig) {
        double sch_v[2], ssx_v[2];
        oc_cplx mygpvar1 = oc_conj(oc_load(aqsmtemp, brow + igp));                 /* :119 */
        oc_cplx matngmatmgp = oc_product(oc_load(aqsntemp, brow + ig), mygpvar1);   /* :121 */
        oracle_pole_terms(wxt, wtilde_array + 2 * (row + ig), I_eps_array + 2 * (row + ig),
                          sch_v, ssx_v);                                            /* :124 */
        /* ssxa[ig] = 0; ssxa[ig] += product(matngmatmgp, ssx)  (:114, :61) */
        oc_cplx ssxa = oc_make(0.0, 0.0), scha = oc_make(0.0, 0.0);
        oc_acc(&ssxa, oc_product(matngmatmgp, oc_make(ssx_v[0], ssx_v[1])));
        oc_acc(&scha, oc_product(matngmatmgp, oc_make(sch_v[0], sch_v[1])));      /* :62 */
        oc_acc(ssxt, ssxa);                                                         /* :125 */
        oc_acc(scht, scha);                                                         /* :126 */
    }
}

/* noflagOCC_solver, gppComplex.cpp:131-150.  Returns scht. */
static oc_cplx oc_noflag_occ(double wxt, const double *wtilde_array, int my_igp, int n1,
                             const double *aqsmtemp, const double *aqsntemp,
                             const double *I_eps_array, int ncouls, int igp)
{
    const long row = (long)my_igp * ncouls;
    const long brow = (long)n1 * ncouls;
    oc_cplx mygpvar1 = oc_conj(oc_load(aqsmtemp, brow + igp));              /* :134 */
    oc_cplx scht_loc = oc_make(0.0, 0.0);
    for (int ig = 0; ig < ncouls; ++ig) {
        oc_cplx wt = oc_load(wtilde_array, row + ig);
        oc_cplx wdiff = oc_dminus(wxt, wt);                                  /* :139 */
        double wdiffr = oc_product(wdiff, oc_conj(wdiff)).re;                /* :140 */
        double rden = 1 / wdiffr;                                            /* :141 */
        oc_cplx delw = oc_mult(oc_product(wt, oc_conj(wdiff)), rden);        /* :143 */
        oc_acc(&scht_loc, oc_product(oc_product(mygpvar1, oc_load(aqsntemp, brow + ig)),
                                     oc_product(delw, oc_load(I_eps_array, row + ig)))); /* :146 */
    }
    return scht_loc;                                                         /* :149 */
}

/* igp mapping and clamp, gppComplex.cpp:342-344. */
static inline int oc_igp(const int *inv_igp_index, int my_igp, int ncouls)
{
    int igp = inv_igp_index[my_igp];
    if (igp >= ncouls) igp = ncouls - 1;
    return igp;
}

/* reduce_achstemp for one band, gppComplex.cpp:66-101 (negative indexes are
 * rejected by oracle_sigma before this runs). */
static void oc_reduce_achstemp(int n1, const int *inv_igp_index, int ncouls,
                               const double *aqsmtemp, const double *aqsntemp,
                               const double *I_eps_array, oc_cplx *achstemp, int ngpown,
                               const double *vcoul)
{
    const long brow = (long)n1 * ncouls;
    for (int my_igp = 0; my_igp < ngpown; my_igp++) {
        oc_cplx schstemp = oc_make(0.0, 0.0);
        int igp = oc_igp(inv_igp_index, my_igp, ncouls);                     /* :77-79 */
        oc_cplx mygpvar1 = oc_conj(oc_load(aqsmtemp, brow + igp));          /* :84 */
        oc_cplx schs = oc_neg(oc_load(I_eps_array, (long)my_igp * ncouls + igp)); /* :86-88 */
        oc_cplx matngmatmgp = oc_product(oc_load(aqsntemp, brow + igp), mygpvar1); /* :89 */
        if (oc_abs(schs) > OC_TO1) {                                          /* :91-92 */
            oc_cplx p = oc_product(matngmatmgp, schs);
            schstemp.re += p.re;
            schstemp.im += p.im;
        }
        oc_acc(achstemp, oc_mult(schstemp, vcoul[igp] * (double)0.5));        /* :99 */
    }
}

/* Index of the frequency whose sch feeds acht_n1_loc (gppComplex.cpp:381
 * hard-codes sch_array[2]); for nfreq < 3 the last frequency is used. */
static inline int oc_acht_iw(int nfreq) { return nfreq > 2 ? 2 : nfreq - 1; }

static int oc_check(int nb, int nv, int ncouls, int ngpown, int nfreq, const int *inv_igp_index)
{
    if (nb < 0 || nv < 0 || nv > nb || ncouls <= 0 || ngpown < 0 || nfreq <= 0) return 1;
    for (int j = 0; j < ngpown; ++j)
        if (inv_igp_index[j] < 0) return 2;
    return 0;
}

/*
 * The timed band loop of gppComplex.cpp:335-383 (main), with nfreq
 * frequencies instead of the fixed nstart=0/nend=3.  nthreads <= 1 runs the
 * loop sequentially; nthreads > 1 splits each band's my_igp loop over
 * OpenMP threads as gppOpenMP3.cpp:283 does, but race-free: each my_igp
 * writes its per-frequency results to its own slot and the slots are then
 * accumulated in my_igp order, so the result is bitwise equal to the
 * sequential loop.
 *
 * Outputs (interleaved complex, overwritten): achtemp[nfreq], asxtemp[nfreq],
 * acht_n1_loc[nb] (may be NULL), achstemp[1] (may be NULL).
 * Returns 0, 1 for bad sizes, 2 for a negative inv_igp_index entry.
 */
int oracle_sigma(int nb, int nv, int ncouls, int ngpown, int nfreq, double occ,
                 const int *inv_igp_index, const double *wx_array, const double *vcoul,
                 const double *aqsmtemp, const double *aqsntemp,
                 const double *I_eps_array, const double *wtilde_array,
                 double *achtemp_out, double *asxtemp_out, double *acht_n1_loc_out,
                 double *achstemp_out, int nthreads)
{
    return oracle_sigma_wxb(nb, nv, ncouls, ngpown, nfreq, occ, inv_igp_index, wx_array, 0, vcoul, aqsmtemp,
                            aqsntemp, I_eps_array, wtilde_array, achtemp_out, asxtemp_out, acht_n1_loc_out,
                            achstemp_out, nthreads);
}

/*
 * The same loop with band-dependent frequencies: band n1 uses
 * wx[n1 * wx_stride + iw] (wx_stride 0: the reference's shared wx_array;
 * wx_stride nfreq: one row per band, as BerkeleyGW's wx = e_lk - e_n1kq(n1)
 * + dw*(iw-1), which the mini-app fixes to one e_n1kq, gppComplex.cpp:326).
 * Only the wxt of :354 changes; everything else is the loop above.
 */
int oracle_sigma_wxb(int nb, int nv, int ncouls, int ngpown, int nfreq, double occ,
                     const int *inv_igp_index, const double *wx_array, int wx_stride, const double *vcoul,
                     const double *aqsmtemp, const double *aqsntemp,
                     const double *I_eps_array, const double *wtilde_array,
                     double *achtemp_out, double *asxtemp_out, double *acht_n1_loc_out,
                     double *achstemp_out, int nthreads)
{
    int rc = oc_check(nb, nv, ncouls, ngpown, nfreq, inv_igp_index);
    if (rc) return rc;
    if (wx_stride != 0 && wx_stride != nfreq) return 1;
    const int acht_iw = oc_acht_iw(nfreq);
    oc_cplx *achtemp = calloc((size_t)nfreq, sizeof(oc_cplx));
    oc_cplx *asxtemp = calloc((size_t)nfreq, sizeof(oc_cplx));
    oc_cplx achstemp = oc_make(0.0, 0.0);
    /* per-my_igp slots: sch_array[iw] and ssx_array[iw] (gppComplex.cpp:280-281) */
    oc_cplx *sch_slots = calloc((size_t)ngpown * nfreq + 1, sizeof(oc_cplx));
    oc_cplx *ssx_slots = calloc((size_t)ngpown * nfreq + 1, sizeof(oc_cplx));
    if (!achtemp || !asxtemp || !sch_slots || !ssx_slots) {
        free(achtemp); free(asxtemp); free(sch_slots); free(ssx_slots);
        return 3;
    }
    if (nthreads < 1) nthreads = 1;

    for (int n1 = 0; n1 < nb; ++n1) {                                        /* :335 */
        const int flag_occ = n1 < nv;                                        /* :337 */
        oc_reduce_achstemp(n1, inv_igp_index, ncouls, aqsmtemp, aqsntemp, I_eps_array,
                           &achstemp, ngpown, vcoul);                        /* :338 */
        oc_cplx acht_n1 = oc_make(0.0, 0.0);

#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic) num_threads(nthreads) if (nthreads > 1)
#endif
        for (int my_igp = 0; my_igp < ngpown; ++my_igp) {                   /* :340 */
            int igp = oc_igp(inv_igp_index, my_igp, ncouls);                 /* :342-344 */
            oc_cplx *sch_array = sch_slots + (size_t)my_igp * nfreq;
            oc_cplx *ssx_array = ssx_slots + (size_t)my_igp * nfreq;
            for (int i = 0; i < nfreq; i++) {                                /* :347-351 */
                ssx_array[i] = oc_make(0.0, 0.0);
                sch_array[i] = oc_make(0.0, 0.0);
            }
            for (int iw = 0; iw < nfreq; ++iw) {
                double wxt = wx_array[(size_t)n1 * wx_stride + iw];
                if (flag_occ) {                                              /* :355-364 */
                    oc_cplx scht = oc_make(0.0, 0.0), ssxt = oc_make(0.0, 0.0);
                    oc_flag_occ(wxt, wtilde_array, my_igp, n1, aqsmtemp, aqsntemp, I_eps_array,
                                &ssxt, &scht, ncouls, igp);
                    oc_acc(&ssx_array[iw], ssxt);
                    oc_acc(&sch_array[iw], oc_mult(scht, 0.5));
                } else {                                                     /* :368-375 */
                    oc_cplx scht = oc_noflag_occ(wxt, wtilde_array, my_igp, n1, aqsmtemp,
                                                 aqsntemp, I_eps_array, ncouls, igp);
                    oc_acc(&sch_array[iw], oc_mult(scht, 0.5));
                }
            }
        }
        /* Accumulation in my_igp order (:363, :378-381). */
        for (int my_igp = 0; my_igp < ngpown; ++my_igp) {
            int igp = oc_igp(inv_igp_index, my_igp, ncouls);
            const oc_cplx *sch_array = sch_slots + (size_t)my_igp * nfreq;
            const oc_cplx *ssx_array = ssx_slots + (size_t)my_igp * nfreq;
            if (flag_occ)
                for (int iw = 0; iw < nfreq; ++iw)
                    oc_acc(&asxtemp[iw], oc_mult(ssx_array[iw], occ * vcoul[igp]));   /* :363 */
            for (int iw = 0; iw < nfreq; ++iw)
                oc_acc(&achtemp[iw], oc_mult(sch_array[iw], vcoul[igp]));             /* :379 */
            oc_acc(&acht_n1, oc_mult(sch_array[acht_iw], vcoul[igp]));                /* :381 */
        }
        if (acht_n1_loc_out) {
            acht_n1_loc_out[2 * n1] = acht_n1.re;
            acht_n1_loc_out[2 * n1 + 1] = acht_n1.im;
        }
    }
    for (int iw = 0; iw < nfreq; ++iw) {
        achtemp_out[2 * iw] = achtemp[iw].re;
        achtemp_out[2 * iw + 1] = achtemp[iw].im;
        asxtemp_out[2 * iw] = asxtemp[iw].re;
        asxtemp_out[2 * iw + 1] = asxtemp[iw].im;
    }
    if (achstemp_out) { achstemp_out[0] = achstemp.re; achstemp_out[1] = achstemp.im; }
    free(achtemp); free(asxtemp); free(sch_slots); free(ssx_slots);
    return 0;
}

/* ---------------------------------------------------------------------------
 * Compensated variant: the same terms, each computed exactly as above, but
 * every running sum is a double-double (Knuth TwoSum), so the result is the
 * sum of the reference's rounded terms to ~2^-100 relative.  It states how
 * far the reference's own left-to-right sums are from their exact value.
 * ------------------------------------------------------------------------- */
typedef struct { double hi, lo; } oc_dd;
typedef struct { oc_dd re, im; } oc_ddc;

static inline oc_dd oc_dd_add(oc_dd a, double b)
{
    double s = a.hi + b;
    double bb = s - a.hi;
    double err = (a.hi - (s - bb)) + (b - bb);
    double lo = a.lo + err;
    double hi = s + lo;
    oc_dd r; r.lo = lo - (hi - s); r.hi = hi;
    return r;
}
static inline oc_dd oc_dd_add_dd(oc_dd a, oc_dd b) { return oc_dd_add(oc_dd_add(a, b.hi), b.lo); }
static inline void oc_ddc_acc(oc_ddc *a, oc_cplx b) { a->re = oc_dd_add(a->re, b.re); a->im = oc_dd_add(a->im, b.im); }
static inline void oc_ddc_acc_dd(oc_ddc *a, oc_ddc b) { a->re = oc_dd_add_dd(a->re, b.re); a->im = oc_dd_add_dd(a->im, b.im); }
/* (dd complex) * double, to double-double accuracy using fma splitting */
static inline oc_dd oc_dd_mul_d(oc_dd a, double b)
{
    double p = a.hi * b;
    double e = fma(a.hi, b, -p);
    e += a.lo * b;
    double hi = p + e;
    oc_dd r; r.hi = hi; r.lo = e - (hi - p);
    return r;
}
static inline oc_ddc oc_ddc_mul_d(oc_ddc a, double b) { oc_ddc r; r.re = oc_dd_mul_d(a.re, b); r.im = oc_dd_mul_d(a.im, b); return r; }

/* Test hooks for the double-double helpers above (the tests compare them with
 * exact rational sums and products): out = {hi, lo}. */
void oracle_dd_sum(const double *x, long n, double *out)
{
    oc_dd a = {0.0, 0.0};
    for (long i = 0; i < n; ++i) a = oc_dd_add(a, x[i]);
    out[0] = a.hi; out[1] = a.lo;
}
void oracle_dd_sum_dd(const double *hilo, long n, double *out)
{
    oc_dd a = {0.0, 0.0};
    for (long i = 0; i < n; ++i) { oc_dd b = {hilo[2 * i], hilo[2 * i + 1]}; a = oc_dd_add_dd(a, b); }
    out[0] = a.hi; out[1] = a.lo;
}
void oracle_dd_mul_d(double hi, double lo, double b, double *out)
{
    oc_dd a = {hi, lo};
    oc_dd r = oc_dd_mul_d(a, b);
    out[0] = r.hi; out[1] = r.lo;
}

int oracle_sigma_dd2(int nb, int nv, int ncouls, int ngpown, int nfreq, double occ,
                     const int *inv_igp_index, const double *wx_array, const double *vcoul,
                     const double *aqsmtemp, const double *aqsntemp,
                     const double *I_eps_array, const double *wtilde_array,
                     double *achtemp_out, double *asxtemp_out, double *acht_n1_loc_out,
                     double *achstemp_out, int nthreads);

int oracle_sigma_dd(int nb, int nv, int ncouls, int ngpown, int nfreq, double occ,
                    const int *inv_igp_index, const double *wx_array, const double *vcoul,
                    const double *aqsmtemp, const double *aqsntemp,
                    const double *I_eps_array, const double *wtilde_array,
                    double *achtemp_out, double *asxtemp_out, int nthreads)
{
    return oracle_sigma_dd2(nb, nv, ncouls, ngpown, nfreq, occ, inv_igp_index, wx_array, vcoul, aqsmtemp,
                            aqsntemp, I_eps_array, wtilde_array, achtemp_out, asxtemp_out, NULL, NULL, nthreads);
}

/* The compensated loop with every output: acht_n1_loc[nb] (gppComplex.cpp:381,
 * each band's frequency-acht_iw terms summed in double-double) and achstemp
 * (reduce_achstemp, :66-101, its per-column terms summed in double-double);
 * either may be NULL. */
int oracle_sigma_dd2(int nb, int nv, int ncouls, int ngpown, int nfreq, double occ,
                     const int *inv_igp_index, const double *wx_array, const double *vcoul,
                     const double *aqsmtemp, const double *aqsntemp,
                     const double *I_eps_array, const double *wtilde_array,
                     double *achtemp_out, double *asxtemp_out, double *acht_n1_loc_out,
                     double *achstemp_out, int nthreads)
{
    int rc = oc_check(nb, nv, ncouls, ngpown, nfreq, inv_igp_index);
    if (rc) return rc;
    const long nslots = (long)nb * ngpown * nfreq;
    oc_ddc *ach_slots = calloc((size_t)nslots + 1, sizeof(oc_ddc));
    oc_ddc *asx_slots = calloc((size_t)nslots + 1, sizeof(oc_ddc));
    if (!ach_slots || !asx_slots) { free(ach_slots); free(asx_slots); return 3; }
    if (nthreads < 1) nthreads = 1;
    const long npairs = (long)nb * ngpown;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic) num_threads(nthreads) if (nthreads > 1)
#endif
    for (long p = 0; p < npairs; ++p) {
        const int n1 = (int)(p / ngpown), my_igp = (int)(p % ngpown);
        const int igp = oc_igp(inv_igp_index, my_igp, ncouls);
        const long row = (long)my_igp * ncouls, brow = (long)n1 * ncouls;
        const oc_cplx mygpvar1 = oc_conj(oc_load(aqsmtemp, brow + igp));
        for (int iw = 0; iw < nfreq; ++iw) {
            const double wxt = wx_array[iw];
            oc_ddc scht, ssxt;
            memset(&scht, 0, sizeof scht); memset(&ssxt, 0, sizeof ssxt);
            for (int ig = 0; ig < ncouls; ++ig) {
                if (n1 < nv) {
                    double sch_v[2], ssx_v[2];
                    oc_cplx m = oc_product(oc_load(aqsntemp, brow + ig), mygpvar1);
                    oracle_pole_terms(wxt, wtilde_array + 2 * (row + ig), I_eps_array + 2 * (row + ig), sch_v, ssx_v);
                    oc_ddc_acc(&ssxt, oc_product(m, oc_make(ssx_v[0], ssx_v[1])));
                    oc_ddc_acc(&scht, oc_product(m, oc_make(sch_v[0], sch_v[1])));
                } else {
                    oc_cplx wt = oc_load(wtilde_array, row + ig);
                    oc_cplx wdiff = oc_dminus(wxt, wt);
                    double rden = 1 / oc_product(wdiff, oc_conj(wdiff)).re;
                    oc_cplx delw = oc_mult(oc_product(wt, oc_conj(wdiff)), rden);
                    oc_ddc_acc(&scht, oc_product(oc_product(mygpvar1, oc_load(aqsntemp, brow + ig)),
                                                 oc_product(delw, oc_load(I_eps_array, row + ig))));
                }
            }
            ach_slots[p * nfreq + iw] = oc_ddc_mul_d(oc_ddc_mul_d(scht, 0.5), vcoul[igp]);
            asx_slots[p * nfreq + iw] = oc_ddc_mul_d(ssxt, occ * vcoul[igp]);
        }
    }
    for (int iw = 0; iw < nfreq; ++iw) {
        oc_ddc a, s;
        memset(&a, 0, sizeof a); memset(&s, 0, sizeof s);
        for (long p = 0; p < npairs; ++p) {
            oc_ddc_acc_dd(&a, ach_slots[p * nfreq + iw]);
            if (p / ngpown < nv) oc_ddc_acc_dd(&s, asx_slots[p * nfreq + iw]);
        }
        achtemp_out[2 * iw] = a.re.hi + a.re.lo;
        achtemp_out[2 * iw + 1] = a.im.hi + a.im.lo;
        asxtemp_out[2 * iw] = s.re.hi + s.re.lo;
        asxtemp_out[2 * iw + 1] = s.im.hi + s.im.lo;
    }
    if (acht_n1_loc_out) {
        const int acht_iw = oc_acht_iw(nfreq);
        for (int n1 = 0; n1 < nb; ++n1) {
            oc_ddc a;
            memset(&a, 0, sizeof a);
            for (int j = 0; j < ngpown; ++j) oc_ddc_acc_dd(&a, ach_slots[((long)n1 * ngpown + j) * nfreq + acht_iw]);
            acht_n1_loc_out[2 * n1] = a.re.hi + a.re.lo;
            acht_n1_loc_out[2 * n1 + 1] = a.im.hi + a.im.lo;
        }
    }
    if (achstemp_out) {  /* reduce_achstemp's terms, :84-99, in double-double */
        oc_ddc a;
        memset(&a, 0, sizeof a);
        for (int n1 = 0; n1 < nb; ++n1)
            for (int j = 0; j < ngpown; ++j) {
                const int igp = oc_igp(inv_igp_index, j, ncouls);
                const long brow = (long)n1 * ncouls;
                const oc_cplx mygpvar1 = oc_conj(oc_load(aqsmtemp, brow + igp));
                const oc_cplx schs = oc_neg(oc_load(I_eps_array, (long)j * ncouls + igp));
                const oc_cplx m = oc_product(oc_load(aqsntemp, brow + igp), mygpvar1);
                if (oc_abs(schs) > OC_TO1) {
                    oc_ddc t;
                    memset(&t, 0, sizeof t);
                    oc_ddc_acc(&t, oc_product(m, schs));
                    oc_ddc_acc_dd(&a, oc_ddc_mul_d(t, vcoul[igp] * (double)0.5));
                }
            }
        achstemp_out[0] = a.re.hi + a.re.lo;
        achstemp_out[1] = a.im.hi + a.im.lo;
    }
    free(ach_slots); free(asx_slots);
    return 0;
}

/* Threads OpenMP would use by default (1 when built without OpenMP). */
int oracle_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
