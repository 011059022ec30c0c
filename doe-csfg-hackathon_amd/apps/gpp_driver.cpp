// gpp_driver -- command-line driver of the GPP sigma mini-app on MI355X.
//
// Keeps the reference mini-app's interface (gppComplex.cpp:219-405): the four
// positional arguments <number_bands> <nvband> <ncouls> <nodes_per_group>,
// ngpown = ncouls / nodes_per_group (:238), the parameter banner (:250-260),
// the array-size lines (:287-289), the deterministic initializer (:296-329),
// the achtemp printout "( %f, %f) " (:390-391, Complex.h:76-79) and the
// "Time Taken" line (:393).  The timed band loop (:332-387) is replaced by
// one call into libgpp_hip.so (include/gpp.h).  Unlike the reference, a
// usage error exits with status 2, not 0 (:223-228).
//
// Extra options:
//   --form direct|direct-strict|factored-f64|factored-ozaki  (default direct)
//   --nfreq N --dw D --wx-center C   frequency grid (seeded init)
//   --init det|seeded --seed S       input generator (default det)
//   --igp-index plus1|zero           det init index map (default plus1 =
//                                    reference (i+1)*ncouls/ngpown)
//   --ngpown N                       override ncouls / nodes_per_group
//   --device D --repeat R --json     device, timed repetitions, JSON line
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gpp.h"

namespace {

void usage()
{
    std::printf("The correct form of input is : \n");
    std::printf(" ./gpp_driver <number_bands> <number_valence_bands> <number_plane_waves> <matrix_divider> "
                "[--form direct|direct-strict|factored-f64|factored-ozaki] [--init det|seeded] [--seed S] "
                "[--nfreq N] [--dw D] [--wx-center C] [--igp-index plus1|zero] [--ngpown N] [--device D] "
                "[--repeat R] [--json]\n");
}

bool parse_form(const std::string &s, gpp_form *f)
{
    if (s == "direct") *f = GPP_FORM_DIRECT;
    else if (s == "direct-strict") *f = GPP_FORM_DIRECT_STRICT;
    else if (s == "factored-f64") *f = GPP_FORM_FACTORED_F64;
    else if (s == "factored-ozaki") *f = GPP_FORM_FACTORED_OZAKI;
    else return false;
    return true;
}

int die(const char *what, gpp_status st)
{
    std::fprintf(stderr, "%s failed (status %d): %s\n", what, (int)st, gpp_last_error());
    return 1;
}

}  // namespace

int main(int argc, char **argv)
{
    std::vector<std::string> pos;
    gpp_form form = GPP_FORM_DIRECT;
    std::string form_name = "direct", init = "det";
    unsigned long long seed = 1;
    int nfreq = 3, device = 0, ngpown_override = 0, repeat = 1;
    double dw = 1.0, wx_center = 4.0;
    bool igp_zero = false, json = false;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&](void) -> const char * {
            if (i + 1 >= argc) {
                usage();
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "--form") {
            form_name = next();
            if (!parse_form(form_name, &form)) { usage(); return 2; }
        } else if (a == "--init") init = next();
        else if (a == "--seed") seed = std::strtoull(next(), nullptr, 10);
        else if (a == "--nfreq") nfreq = std::atoi(next());
        else if (a == "--dw") dw = std::atof(next());
        else if (a == "--wx-center") wx_center = std::atof(next());
        else if (a == "--igp-index") igp_zero = std::string(next()) == "zero";
        else if (a == "--ngpown") ngpown_override = std::atoi(next());
        else if (a == "--device") device = std::atoi(next());
        else if (a == "--repeat") repeat = std::atoi(next());
        else if (a == "--json") json = true;
        else if (a.size() > 1 && a[0] == '-' && !(a[1] >= '0' && a[1] <= '9')) { usage(); return 2; }
        else pos.push_back(a);
    }
    if (pos.size() != 4 || (init != "det" && init != "seeded") || repeat < 1) {
        usage();
        return 2;
    }
    const int number_bands = std::atoi(pos[0].c_str());
    const int nvband = std::atoi(pos[1].c_str());
    const int ncouls = std::atoi(pos[2].c_str());
    const int nodes_per_group = std::atoi(pos[3].c_str());
    if (nodes_per_group <= 0) { usage(); return 2; }
    const int ngpown = ngpown_override > 0 ? ngpown_override : ncouls / nodes_per_group;

    // Banner, gppComplex.cpp:250-260.
    std::printf("number_bands = %d\t nvband = %d\t ncouls = %d\t nodes_per_group  = %d\t ngpown = %d\t nend = %d\t "
                "nstart = %d\t gamma = %g\t sexcut = %g\t limitone = %g\t limittwo = %g\n",
                number_bands, nvband, ncouls, nodes_per_group, ngpown, nfreq, 0, 0.5, 4.0, 1.0 / (1e-6 * 4.0), 0.25);
    std::printf("Size of wtilde_array = %g Mbytes\n", (ncouls * (double)ngpown * 2.0 * 8) / std::pow(1024, 2));
    std::printf("Size of aqsntemp = %g Mbytes\n", (ncouls * (double)number_bands * 2.0 * 8) / std::pow(1024, 2));
    std::printf("Size of I_eps_array array = %g Mbytes\n", (ncouls * (double)ngpown * 2.0 * 8) / std::pow(1024, 2));

    gpp_problem prob{number_bands, nvband, ncouls, ngpown, nfreq, 1.0};
    gpp_ctx *ctx = nullptr;
    gpp_status st = gpp_ctx_create(&ctx, device);
    if (st) return die("gpp_ctx_create", st);

    if (init == "det") {
        std::vector<int32_t> idx(ngpown > 0 ? ngpown : 1);
        std::vector<double> wx(nfreq > 0 ? nfreq : 1), vc(ncouls > 0 ? ncouls : 1);
        std::vector<gpp_complex> aqs((size_t)std::max(number_bands, 1) * std::max(ncouls, 1));
        std::vector<gpp_complex> eps((size_t)std::max(ngpown, 1) * std::max(ncouls, 1));
        st = gpp_fill_deterministic(&prob, igp_zero ? 1 : 0, idx.data(), wx.data(), vc.data(), aqs.data(), nullptr,
                                    eps.data(), nullptr);
        if (st) return die("gpp_fill_deterministic", st);
        // aqsmtemp == aqsntemp and I_eps_array == wtilde_array at this init (:299-309)
        st = gpp_upload(ctx, &prob, idx.data(), wx.data(), vc.data(), aqs.data(), aqs.data(), eps.data(), eps.data());
        if (st) return die("gpp_upload", st);
    } else {
        st = gpp_generate_seeded(ctx, &prob, seed, wx_center, dw);
        if (st) return die("gpp_generate_seeded", st);
    }

    std::vector<gpp_complex> achtemp(nfreq), asxtemp(nfreq), acht_n1((size_t)number_bands);
    gpp_complex achstemp{0, 0};
    gpp_outputs out{achtemp.data(), asxtemp.data(), acht_n1.data(), &achstemp};
    double best_gpu_ms = 1e300, wall_s = 0;
    for (int r = 0; r < repeat; ++r) {
        auto t0 = std::chrono::steady_clock::now();
        st = gpp_sigma(ctx, form, &out);
        auto t1 = std::chrono::steady_clock::now();
        if (st) return die("gpp_sigma", st);
        double ms = 0;
        gpp_ctx_last_timing(ctx, &ms, nullptr, 0, nullptr);
        if (ms < best_gpu_ms) best_gpu_ms = ms;
        wall_s = std::chrono::duration<double>(t1 - t0).count();
    }

    for (int iw = 0; iw < nfreq; ++iw) std::printf("( %f, %f) \n", achtemp[iw].re, achtemp[iw].im);
    std::printf("********** Time Taken **********= %g secs\n", wall_s);
    if (json) {
        const double updates = (double)number_bands * ngpown * ncouls * nfreq;
        std::printf("{\"form\": \"%s\", \"number_bands\": %d, \"nvband\": %d, \"ncouls\": %d, \"ngpown\": %d, "
                    "\"nfreq\": %d, \"gpu_ms\": %.6f, \"updates_per_s\": %.6e, \"achtemp\": [",
                    form_name.c_str(), number_bands, nvband, ncouls, ngpown, nfreq, best_gpu_ms,
                    updates / (best_gpu_ms * 1e-3));
        for (int iw = 0; iw < nfreq; ++iw)
            std::printf("%s[%.17g, %.17g]", iw ? ", " : "", achtemp[iw].re, achtemp[iw].im);
        std::printf("], \"asxtemp\": [");
        for (int iw = 0; iw < nfreq; ++iw)
            std::printf("%s[%.17g, %.17g]", iw ? ", " : "", asxtemp[iw].re, asxtemp[iw].im);
        std::printf("], \"achstemp\": [%.17g, %.17g]}\n", achstemp.re, achstemp.im);
    }
    gpp_ctx_destroy(ctx);
    return 0;
}
