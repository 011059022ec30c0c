/*
 * tsan_check.c -- TEST INFRASTRUCTURE ONLY: drives the oracle's OpenMP loops
 * under ThreadSanitizer (`make -C oracle sanitize`).
 *
 * The reference's OpenMP build shares ssxa/scha scratch and accumulates
 * achtemp from every thread (gppOpenMP3.cpp:306, :310); the oracle's OpenMP
 * restatement gives each my_igp its own result slots and sums them in order.
 * This program runs oracle_sigma and oracle_sigma_dd with several threads on
 * a small seeded problem (occupied and unoccupied bands, several
 * frequencies) and checks that the threaded results equal the sequential
 * ones bit for bit; TSan reports any data race in the loops.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int oracle_sigma(int nb, int nv, int ncouls, int ngpown, int nfreq, double occ, const int *inv_igp_index,
                 const double *wx_array, const double *vcoul, const double *aqsmtemp, const double *aqsntemp,
                 const double *I_eps_array, const double *wtilde_array, double *achtemp_out, double *asxtemp_out,
                 double *acht_n1_loc_out, double *achstemp_out, int nthreads);
int oracle_sigma_dd(int nb, int nv, int ncouls, int ngpown, int nfreq, double occ, const int *inv_igp_index,
                    const double *wx_array, const double *vcoul, const double *aqsmtemp, const double *aqsntemp,
                    const double *I_eps_array, const double *wtilde_array, double *achtemp_out,
                    double *asxtemp_out, int nthreads);

static uint64_t state = 0x9E3779B97F4A7C15ull;
static double uni(double lo, double hi)
{
    state = state * 6364136223846793005ull + 1442695040888963407ull;
    return lo + (hi - lo) * (double)(state >> 11) * (1.0 / 9007199254740992.0);
}

int main(void)
{
    enum { NB = 10, NV = 3, NC = 96, NG = 24, NF = 4 };
    static double aqsm[2 * NB * NC], aqsn[2 * NB * NC], ieps[2 * NG * NC], wt[2 * NG * NC], vc[NC];
    static int idx[NG];
    double wx[NF] = {2.5, 3.5, 4.5, 5.5};
    for (int i = 0; i < 2 * NB * NC; ++i) { aqsm[i] = uni(-1, 1); aqsn[i] = uni(-1, 1); }
    for (int i = 0; i < NG * NC; ++i) {
        ieps[2 * i] = uni(-0.5, 0.5); ieps[2 * i + 1] = uni(-0.5, 0.5);
        wt[2 * i] = uni(0.5, 5.0);    wt[2 * i + 1] = uni(-0.05, 0.05);
    }
    for (int i = 0; i < NC; ++i) vc[i] = uni(0.1, 10);
    for (int j = 0; j < NG; ++j) idx[j] = (j * 37) % NC;
    idx[NG - 1] = NC + 5;  /* clamped to ncouls - 1 */

    double a1[2 * NF], x1[2 * NF], n1[2 * NB], h1[2], a4[2 * NF], x4[2 * NF], n4[2 * NB], h4[2];
    if (oracle_sigma(NB, NV, NC, NG, NF, 1.0, idx, wx, vc, aqsm, aqsn, ieps, wt, a1, x1, n1, h1, 1) ||
        oracle_sigma(NB, NV, NC, NG, NF, 1.0, idx, wx, vc, aqsm, aqsn, ieps, wt, a4, x4, n4, h4, 4)) {
        fprintf(stderr, "oracle_sigma failed\n");
        return 1;
    }
    if (memcmp(a1, a4, sizeof a1) || memcmp(x1, x4, sizeof x1) || memcmp(n1, n4, sizeof n1) ||
        memcmp(h1, h4, sizeof h1)) {
        fprintf(stderr, "threaded oracle_sigma differs from the sequential loop\n");
        return 1;
    }
    double d1[2 * NF], e1[2 * NF], d4[2 * NF], e4[2 * NF];
    if (oracle_sigma_dd(NB, NV, NC, NG, NF, 1.0, idx, wx, vc, aqsm, aqsn, ieps, wt, d1, e1, 1) ||
        oracle_sigma_dd(NB, NV, NC, NG, NF, 1.0, idx, wx, vc, aqsm, aqsn, ieps, wt, d4, e4, 4)) {
        fprintf(stderr, "oracle_sigma_dd failed\n");
        return 1;
    }
    if (memcmp(d1, d4, sizeof d1) || memcmp(e1, e4, sizeof e1)) {
        fprintf(stderr, "threaded oracle_sigma_dd differs from one thread\n");
        return 1;
    }
    printf("tsan_check: oracle_sigma and oracle_sigma_dd on 4 threads equal one thread bit for bit\n");
    return 0;
}
