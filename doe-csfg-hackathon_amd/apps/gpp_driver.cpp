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
//   --load FILE                      inputs from a problem file (.gpp, the
//                                    positional arguments are then optional);
//                                    reports the error against the file's
//                                    expected outputs when it holds them
//   --save FILE                      writes the inputs and the computed
//                                    outputs (as expected outputs) to FILE
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
                "[--form direct|direct-strict|factored-f64|factored-ozaki|factored-ozaki-i8] [--init det|seeded] "
                "[--seed S] [--nfreq N] [--dw D] [--wx-center C] [--igp-index plus1|zero] [--ngpown N] "
                "[--device D] [--repeat R] [--json] [--save FILE]\n");
    std::printf(" ./gpp_driver --load FILE [--form ...] [--device D] [--repeat R] [--json] [--save FILE]\n");
}

bool parse_form(const std::string &s, gpp_form *f)
{
    if (s == "direct") *f = GPP_FORM_DIRECT;
    else if (s == "direct-strict") *f = GPP_FORM_DIRECT_STRICT;
    else if (s == "factored-f64") *f = GPP_FORM_FACTORED_F64;
    else if (s == "factored-ozaki") *f = GPP_FORM_FACTORED_OZAKI;
    else if (s == "factored-ozaki-i8") *f = GPP_FORM_FACTORED_OZAKI_I8;
    else return false;
    return true;
}

int die(const char *what, gpp_status st)
{
    std::fprintf(stderr, "%s failed (status %d): %s\n", what, (int)st, gpp_last_error());
    return 1;
}

// Largest per-component relative error of got against ref (components below
// 1e-6 of |ref| are compared relative to |ref|); NaN references are skipped.
double rel_err(const gpp_complex *got, const gpp_complex *ref, size_t n)
{
    double worst = 0;
    for (size_t i = 0; i < n; ++i) {
        const double m = std::hypot(ref[i].re, ref[i].im);
        if (std::isnan(m)) continue;
        const double pr[2][2] = {{got[i].re, ref[i].re}, {got[i].im, ref[i].im}};
        for (auto &c : pr) {
            const double den = std::max({std::fabs(c[1]), 1e-6 * m, 1e-300});
            worst = std::max(worst, std::fabs(c[0] - c[1]) / den);
        }
    }
    return worst;
}

}  // namespace

int main(int argc, char **argv)
{
    std::vector<std::string> pos;
    gpp_form form = GPP_FORM_DIRECT;
    std::string form_name = "direct", init = "det", load_path, save_path;
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
        else if (a == "--load") load_path = next();
        else if (a == "--save") save_path = next();
        else if (a.size() > 1 && a[0] == '-' && !(a[1] >= '0' && a[1] <= '9')) { usage(); return 2; }
        else pos.push_back(a);
    }
    const bool from_file = !load_path.empty();
    if ((!from_file && pos.size() != 4) || (from_file && !pos.empty()) || (init != "det" && init != "seeded") ||
        repeat < 1) {
        usage();
        return 2;
    }
    gpp_file_header fh{};
    if (from_file) {
        gpp_status fst = gpp_file_info(load_path.c_str(), &fh);
        if (fst) return die("gpp_file_info", fst);
        nfreq = fh.problem.nfreq;
    }
    const int number_bands = from_file ? fh.problem.number_bands : std::atoi(pos[0].c_str());
    const int nvband = from_file ? fh.problem.nvband : std::atoi(pos[1].c_str());
    const int ncouls = from_file ? fh.problem.ncouls : std::atoi(pos[2].c_str());
    const int nodes_per_group = from_file ? 0 : std::atoi(pos[3].c_str());
    if (!from_file && nodes_per_group <= 0) { usage(); return 2; }
    const int ngpown = from_file ? fh.problem.ngpown
                                 : (ngpown_override > 0 ? ngpown_override : ncouls / nodes_per_group);

    // Banner, gppComplex.cpp:250-260.
    std::printf("number_bands = %d\t nvband = %d\t ncouls = %d\t nodes_per_group  = %d\t ngpown = %d\t nend = %d\t "
                "nstart = %d\t gamma = %g\t sexcut = %g\t limitone = %g\t limittwo = %g\n",
                number_bands, nvband, ncouls, nodes_per_group, ngpown, nfreq, 0, 0.5, 4.0, 1.0 / (1e-6 * 4.0), 0.25);
    std::printf("Size of wtilde_array = %g Mbytes\n", (ncouls * (double)ngpown * 2.0 * 8) / std::pow(1024, 2));
    std::printf("Size of aqsntemp = %g Mbytes\n", (ncouls * (double)number_bands * 2.0 * 8) / std::pow(1024, 2));
    std::printf("Size of I_eps_array array = %g Mbytes\n", (ncouls * (double)ngpown * 2.0 * 8) / std::pow(1024, 2));

    gpp_problem prob{number_bands, nvband, ncouls, ngpown, nfreq, from_file ? fh.problem.occ : 1.0};
    gpp_ctx *ctx = nullptr;
    gpp_status st = gpp_ctx_create(&ctx, device);
    if (st) return die("gpp_ctx_create", st);

    std::vector<gpp_complex> achtemp(nfreq), asxtemp(nfreq), acht_n1((size_t)number_bands);
    gpp_complex achstemp{0, 0};
    gpp_outputs out{achtemp.data(), asxtemp.data(), acht_n1.data(), &achstemp};
    std::vector<gpp_complex> e_ach(nfreq), e_asx(nfreq), e_acht((size_t)number_bands);
    gpp_complex e_achs{0, 0};
    gpp_outputs expected{e_ach.data(), e_asx.data(), e_acht.data(), &e_achs};

    // Host copies of the inputs, kept only when --save needs them.
    std::vector<int32_t> h_idx;
    std::vector<double> h_wx, h_vc;
    std::vector<gpp_complex> h_aqsm, h_aqsn, h_eps, h_wt;
    const bool keep = !save_path.empty();
    if (from_file) {
        st = gpp_upload_file(ctx, load_path.c_str(), &expected);
        if (st) return die("gpp_upload_file", st);
        if (keep) {
            const size_t nwx = (size_t)nfreq * ((fh.flags & GPP_FILE_WX_PER_BAND) ? number_bands : 1);
            h_idx.resize(ngpown); h_wx.resize(nwx); h_vc.resize(ncouls);
            h_aqsm.resize((size_t)number_bands * ncouls); h_aqsn.resize(h_aqsm.size());
            h_eps.resize((size_t)ngpown * ncouls); h_wt.resize(h_eps.size());
            st = gpp_file_read(load_path.c_str(), h_idx.data(), h_wx.data(), h_vc.data(), h_aqsm.data(),
                               h_aqsn.data(), h_eps.data(), h_wt.data(), nullptr);
            if (st) return die("gpp_file_read", st);
        }
    } else if (init == "det" || keep) {
        h_idx.resize(ngpown); h_wx.resize(nfreq); h_vc.resize(ncouls);
        h_aqsm.resize((size_t)number_bands * ncouls);
        h_eps.resize((size_t)ngpown * ncouls);
        if (init == "det") {
            st = gpp_fill_deterministic(&prob, igp_zero ? 1 : 0, h_idx.data(), h_wx.data(), h_vc.data(),
                                        h_aqsm.data(), nullptr, h_eps.data(), nullptr);
            if (st) return die("gpp_fill_deterministic", st);
            // aqsmtemp == aqsntemp and I_eps_array == wtilde_array at this init (:299-309)
            h_aqsn = h_aqsm;
            h_wt = h_eps;
        } else {
            h_aqsn.resize(h_aqsm.size());
            h_wt.resize(h_eps.size());
            st = gpp_fill_seeded(&prob, seed, wx_center, dw, h_idx.data(), h_wx.data(), h_vc.data(), h_aqsm.data(),
                                 h_aqsn.data(), h_eps.data(), h_wt.data());
            if (st) return die("gpp_fill_seeded", st);
        }
        st = gpp_upload(ctx, &prob, h_idx.data(), h_wx.data(), h_vc.data(), h_aqsm.data(), h_aqsn.data(),
                        h_eps.data(), h_wt.data());
        if (st) return die("gpp_upload", st);
    } else {
        st = gpp_generate_seeded(ctx, &prob, seed, wx_center, dw);
        if (st) return die("gpp_generate_seeded", st);
    }
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
    if (from_file && (fh.flags & GPP_FILE_HAS_EXPECTED))
        std::printf("relative error vs expected: achtemp %.3e asxtemp %.3e acht_n1_loc %.3e achstemp %.3e\n",
                    rel_err(achtemp.data(), e_ach.data(), nfreq), rel_err(asxtemp.data(), e_asx.data(), nfreq),
                    rel_err(acht_n1.data(), e_acht.data(), number_bands), rel_err(&achstemp, &e_achs, 1));
    if (keep) {
        const int per_band = from_file && (fh.flags & GPP_FILE_WX_PER_BAND) ? 1 : 0;
        st = gpp_file_write(save_path.c_str(), &prob, h_idx.data(), h_wx.data(), per_band, h_vc.data(),
                            h_aqsm.data(), h_aqsn.data(), h_eps.data(), h_wt.data(), &out);
        if (st) return die("gpp_file_write", st);
    }
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
