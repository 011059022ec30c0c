// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A small C-callable harness around the reference mini-app's own, unmodified
// solver functions.  The recipe in oracle/Makefile compiles the reference's
// gppComplex.cpp in place (its main renamed with -Dmain=gpp_ref_main) and
// links this file beside it into oracle/_ref/libgppref.so, so that seeded and
// near-pole inputs -- which the reference's own deterministic initializer
// never produces -- can be run through the reference's arithmetic and used
// to pin the oracle.  Nothing here is shipped; only tests and the fixture
// script under tests/golden load it.
//
// The band loop below repeats the call sequence of gppComplex.cpp:335-383
// (main) so the reference functions see exactly the arguments main gives
// them; the arithmetic is all inside the reference's functions.
#include "Complex.h"

// Signatures of the reference's functions, gppComplex.cpp:16, :66, :107, :131.
void ssxt_scht_solver(double wxt, int igp, int my_igp, int ig, GPPComplex wtilde, GPPComplex wtilde2,
                      GPPComplex Omega2, GPPComplex matngmatmgp, GPPComplex matngpmatmg,
                      GPPComplex mygpvar1, GPPComplex mygpvar2, GPPComplex &ssxa, GPPComplex &scha,
                      GPPComplex I_eps_array_igp_myIgp);
void reduce_achstemp(int n1, int *inv_igp_index, int ncouls, GPPComplex *aqsmtemp,
                     GPPComplex *aqsntemp, GPPComplex *I_eps_array, GPPComplex &achstemp, int ngpown,
                     double *vcoul);
void flagOCC_solver(double wxt, GPPComplex *wtilde_array, int my_igp, int n1, GPPComplex *aqsmtemp,
                    GPPComplex *aqsntemp, GPPComplex *I_eps_array, GPPComplex &ssxt,
                    GPPComplex &scht, int ncouls, int igp, GPPComplex *ssxa, GPPComplex *scha);
void noflagOCC_solver(double wxt, GPPComplex *wtilde_array, int my_igp, int n1,
                      GPPComplex *aqsmtemp, GPPComplex *aqsntemp, GPPComplex *I_eps_array,
                      GPPComplex &ssxt, GPPComplex &scht, int ncouls, int igp, GPPComplex *scha);
int gpp_ref_main(int argc, char **argv);

static_assert(sizeof(GPPComplex) == 2 * sizeof(double), "GPPComplex must be two packed doubles");

extern "C" {

// The reference's ssxt_scht_solver for one (wxt, wtilde, I_eps) with the
// matrix element set to 1, which returns its raw ssx and sch.
void ref_pole_terms(double wxt, const double *wtilde, const double *ieps, double *sch_out,
                    double *ssx_out)
{
    GPPComplex wt(wtilde[0], wtilde[1]);
    GPPComplex ie(ieps[0], ieps[1]);
    GPPComplex wt2 = GPPComplex_square(wt);
    GPPComplex omega2 = GPPComplex_product(wt2, ie);
    GPPComplex one(1.0, 0.0), zero(0.0, 0.0);
    GPPComplex ssxa(0.0, 0.0), scha(0.0, 0.0);
    ssxt_scht_solver(wxt, 0, 0, 0, wt, wt2, omega2, one, zero, one, one, ssxa, scha, ie);
    sch_out[0] = GPPComplex_real(scha);
    sch_out[1] = GPPComplex_imag(scha);
    ssx_out[0] = GPPComplex_real(ssxa);
    ssx_out[1] = GPPComplex_imag(ssxa);
}

// The band loop of gppComplex.cpp:335-383 over caller-provided inputs, with
// nfreq frequencies; outputs are overwritten.  acht_n1_loc uses frequency 2
// (frequency nfreq-1 when nfreq < 3), as gppComplex.cpp:381.
int ref_sigma(int nb, int nv, int ncouls, int ngpown, int nfreq, double occ, const int *inv_igp_index,
              const double *wx_array, const double *vcoul, const double *aqsmtemp,
              const double *aqsntemp, const double *I_eps_array, const double *wtilde_array,
              double *achtemp_out, double *asxtemp_out, double *acht_n1_loc_out, double *achstemp_out)
{
    if (nb < 0 || nv < 0 || nv > nb || ncouls <= 0 || ngpown < 0 || nfreq <= 0) return 1;
    for (int j = 0; j < ngpown; ++j)
        if (inv_igp_index[j] < 0) return 2;
    const int acht_iw = nfreq > 2 ? 2 : nfreq - 1;
    int *idx = const_cast<int *>(inv_igp_index);
    double *vc = const_cast<double *>(vcoul);
    GPPComplex *am = reinterpret_cast<GPPComplex *>(const_cast<double *>(aqsmtemp));
    GPPComplex *an = reinterpret_cast<GPPComplex *>(const_cast<double *>(aqsntemp));
    GPPComplex *ie = reinterpret_cast<GPPComplex *>(const_cast<double *>(I_eps_array));
    GPPComplex *wt = reinterpret_cast<GPPComplex *>(const_cast<double *>(wtilde_array));
    GPPComplex *achtemp = new GPPComplex[nfreq];
    GPPComplex *asxtemp = new GPPComplex[nfreq];
    GPPComplex *sch_array = new GPPComplex[nfreq];
    GPPComplex *ssx_array = new GPPComplex[nfreq];
    GPPComplex *ssxa = new GPPComplex[ncouls];
    GPPComplex *scha = new GPPComplex[ncouls];
    GPPComplex achstemp(0.0, 0.0), expr0(0.0, 0.0), scht, ssxt;
    for (int n1 = 0; n1 < nb; ++n1) {
        bool flag_occ = n1 < nv;
        GPPComplex acht_n1(0.0, 0.0);
        reduce_achstemp(n1, idx, ncouls, am, an, ie, achstemp, ngpown, vc);
        for (int my_igp = 0; my_igp < ngpown; ++my_igp) {
            int igp = idx[my_igp];
            if (igp >= ncouls) igp = ncouls - 1;
            for (int i = 0; i < nfreq; i++) {
                ssx_array[i] = expr0;
                sch_array[i] = expr0;
            }
            if (flag_occ) {
                for (int iw = 0; iw < nfreq; iw++) {
                    scht = ssxt = expr0;
                    double wxt = wx_array[iw];
                    flagOCC_solver(wxt, wt, my_igp, n1, am, an, ie, ssxt, scht, ncouls, igp, ssxa, scha);
                    ssx_array[iw] += ssxt;
                    sch_array[iw] += GPPComplex_mult(scht, 0.5);
                    asxtemp[iw] += GPPComplex_mult(ssx_array[iw], occ * vc[igp]);
                }
            } else {
                for (int iw = 0; iw < nfreq; ++iw) {
                    scht = ssxt = expr0;
                    double wxt = wx_array[iw];
                    noflagOCC_solver(wxt, wt, my_igp, n1, am, an, ie, ssxt, scht, ncouls, igp, scha);
                    sch_array[iw] += GPPComplex_mult(scht, 0.5);
                }
            }
            for (int iw = 0; iw < nfreq; ++iw)
                achtemp[iw] += GPPComplex_mult(sch_array[iw], vc[igp]);
            acht_n1 += GPPComplex_mult(sch_array[acht_iw], vc[igp]);
        }
        if (acht_n1_loc_out) {
            acht_n1_loc_out[2 * n1] = GPPComplex_real(acht_n1);
            acht_n1_loc_out[2 * n1 + 1] = GPPComplex_imag(acht_n1);
        }
    }
    for (int iw = 0; iw < nfreq; ++iw) {
        achtemp_out[2 * iw] = GPPComplex_real(achtemp[iw]);
        achtemp_out[2 * iw + 1] = GPPComplex_imag(achtemp[iw]);
        asxtemp_out[2 * iw] = GPPComplex_real(asxtemp[iw]);
        asxtemp_out[2 * iw + 1] = GPPComplex_imag(asxtemp[iw]);
    }
    if (achstemp_out) {
        achstemp_out[0] = GPPComplex_real(achstemp);
        achstemp_out[1] = GPPComplex_imag(achstemp);
    }
    delete[] achtemp; delete[] asxtemp; delete[] sch_array; delete[] ssx_array;
    delete[] ssxa; delete[] scha;
    return 0;
}

}  // extern "C"
