// gpp_host_selftest.cpp -- exercises the host-side code of libgpp_hip.so
// (input generators, problem-file writer / reader / header parser, argument
// checks) through the C ABI without touching a GPU, so that it can run under
// AddressSanitizer and UndefinedBehaviorSanitizer on any machine
// (`make -C doe-csfg-hackathon_amd sanitize`).  Exits 0 when every check
// holds; prints the failed check and exits 1 otherwise.
#include <gpp.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

int g_failed = 0;

void check(bool ok, const char *what)
{
    if (!ok) {
        std::fprintf(stderr, "FAILED: %s (last error: %s)\n", what, gpp_last_error());
        ++g_failed;
    }
}

struct Inputs {
    gpp_problem p;
    std::vector<int32_t> idx;
    std::vector<double> wx, vc;
    std::vector<gpp_complex> aqsm, aqsn, ieps, wt;
    explicit Inputs(gpp_problem q, int wx_rows = 1)
        : p(q), idx(q.ngpown), wx((size_t)q.nfreq * wx_rows), vc(q.ncouls), aqsm((size_t)q.number_bands * q.ncouls),
          aqsn(aqsm.size()), ieps((size_t)q.ngpown * q.ncouls), wt(ieps.size())
    {
    }
};

bool same(const std::vector<gpp_complex> &a, const std::vector<gpp_complex> &b)
{
    return a.size() == b.size() && std::memcmp(a.data(), b.data(), sizeof(gpp_complex) * a.size()) == 0;
}

std::vector<unsigned char> slurp(const std::string &path)
{
    std::vector<unsigned char> b;
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return b;
    unsigned char buf[65536];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) b.insert(b.end(), buf, buf + n);
    std::fclose(f);
    return b;
}

void spit(const std::string &path, const std::vector<unsigned char> &b)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return;
    std::fwrite(b.data(), 1, b.size(), f);
    std::fclose(f);
}

void test_fills()
{
    // gppComplex.cpp:296-329 on a small shape
    gpp_problem p{7, 2, 50, 10, 3, 1.0};
    Inputs in(p);
    check(gpp_fill_deterministic(&p, 0, in.idx.data(), in.wx.data(), in.vc.data(), in.aqsm.data(), in.aqsn.data(),
                                 in.ieps.data(), in.wt.data()) == GPP_OK, "fill_deterministic");
    check(in.aqsm[3 * 50 + 4].re == 0.5 * 7 && in.aqsm[3 * 50 + 4].im == 0.5 * 7, "aqsm = (0.5+0.5i)(n1+ig)");
    check(in.ieps[9 * 50 + 49].re == 0.5 * 58, "I_eps = (0.5+0.5i)(j+ig)");
    check(in.idx[9] == 50 && in.idx[0] == 5, "inv_igp_index = (j+1) ncouls / ngpown");
    check(in.wx[0] == 3 && in.wx[2] == 5 && in.vc[49] == 49, "wx and vcoul");
    // NULL outputs are skipped
    check(gpp_fill_deterministic(&p, 1, in.idx.data(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr) == GPP_OK &&
              in.idx[1] == 5,
          "fill_deterministic zero-based index, NULL arrays skipped");
    // seeded: ranges, sorted distinct index, repeatable
    gpp_problem q{5, 1, 300, 299, 16, 1.0};
    Inputs a(q), b(q);
    check(gpp_fill_seeded(&q, 9, 4.0, 0.25, a.idx.data(), a.wx.data(), a.vc.data(), a.aqsm.data(), a.aqsn.data(),
                          a.ieps.data(), a.wt.data()) == GPP_OK, "fill_seeded");
    check(gpp_fill_seeded(&q, 9, 4.0, 0.25, b.idx.data(), b.wx.data(), b.vc.data(), b.aqsm.data(), b.aqsn.data(),
                          b.ieps.data(), b.wt.data()) == GPP_OK && same(a.wt, b.wt) && same(a.aqsn, b.aqsn),
          "fill_seeded repeatable");
    bool ok = true;
    for (int j = 1; j < q.ngpown; ++j) ok = ok && a.idx[j] > a.idx[j - 1] && a.idx[j] < q.ncouls;
    for (const gpp_complex &w : a.wt) ok = ok && w.re >= 0.5 && w.re <= 5.0 && std::fabs(w.im) <= 0.05;
    check(ok, "seeded index sorted and distinct, wtilde in range");
    // wx clamped at 1e-6 (gppComplex.cpp:328)
    gpp_problem c{2, 1, 10, 3, 5, 1.0};
    Inputs cl(c);
    check(gpp_fill_seeded(&c, 1, 0.0, 1.0, nullptr, cl.wx.data(), nullptr, nullptr, nullptr, nullptr, nullptr) ==
                  GPP_OK &&
              cl.wx[0] == 1e-6 && cl.wx[4] == 2.0,
          "seeded wx clamp");
    // argument checks
    gpp_problem bad{4, 5, 10, 3, 3, 1.0};
    check(gpp_fill_deterministic(&bad, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr) ==
              GPP_ERR_INVALID_ARGUMENT,
          "nvband > number_bands refused");
    check(std::strlen(gpp_last_error()) > 0, "error message set");
    check(gpp_fill_deterministic(nullptr, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr) ==
              GPP_ERR_INVALID_ARGUMENT,
          "NULL problem refused");
    gpp_problem wide{4, 1, 10, 20, 3, 1.0};
    check(gpp_fill_seeded(&wide, 1, 4.0, 1.0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr) ==
              GPP_ERR_INVALID_ARGUMENT,
          "seeded ngpown > ncouls refused");
}

void test_files(const std::string &dir)
{
    gpp_problem p{6, 2, 40, 8, 4, 0.75};
    Inputs in(p, p.number_bands);
    check(gpp_fill_seeded(&p, 3, 4.0, 1.0, in.idx.data(), in.wx.data(), in.vc.data(), in.aqsm.data(), in.aqsn.data(),
                          in.ieps.data(), in.wt.data()) == GPP_OK, "fill for files");
    for (int n1 = 0; n1 < p.number_bands; ++n1)  // a frequency row per band
        for (int iw = 0; iw < p.nfreq; ++iw) in.wx[(size_t)n1 * p.nfreq + iw] = 2.0 + 0.5 * n1 + iw;
    std::vector<gpp_complex> ach(p.nfreq), asx(p.nfreq), n1loc(p.number_bands);
    for (int i = 0; i < p.nfreq; ++i) ach[i] = {1.0 + i, -2.0 * i}, asx[i] = {0.5 * i, 3.0};
    gpp_complex achs{7.0, -7.0};
    gpp_outputs exp_out{ach.data(), asx.data(), nullptr, &achs};  // acht_n1_loc stored as NaN
    const std::string path = dir + "/selftest.gpp";
    check(gpp_file_write(path.c_str(), &p, in.idx.data(), in.wx.data(), 1, in.vc.data(), in.aqsm.data(), in.aqsn.data(),
                         in.ieps.data(), in.wt.data(), &exp_out) == GPP_OK, "file_write");
    gpp_file_header h{};
    check(gpp_file_info(path.c_str(), &h) == GPP_OK && h.problem.number_bands == 6 && h.problem.ngpown == 8 &&
              h.problem.occ == 0.75 && h.flags == (GPP_FILE_WX_PER_BAND | GPP_FILE_HAS_EXPECTED),
          "file_info");
    Inputs out(p, p.number_bands);
    std::vector<gpp_complex> r_ach(p.nfreq), r_asx(p.nfreq), r_n1(p.number_bands);
    gpp_complex r_achs{0, 0};
    gpp_outputs r{r_ach.data(), r_asx.data(), r_n1.data(), &r_achs};
    check(gpp_file_read(path.c_str(), out.idx.data(), out.wx.data(), out.vc.data(), out.aqsm.data(), out.aqsn.data(),
                        out.ieps.data(), out.wt.data(), &r) == GPP_OK, "file_read");
    check(out.idx == in.idx && out.wx == in.wx && out.vc == in.vc && same(out.aqsm, in.aqsm) &&
              same(out.aqsn, in.aqsn) && same(out.ieps, in.ieps) && same(out.wt, in.wt),
          "file round trip: inputs");
    check(same(r_ach, ach) && same(r_asx, asx) && r_achs.re == 7.0 && std::isnan(r_n1[0].re),
          "file round trip: expected outputs");
    // damage: truncation, a flipped payload bit, bad magic, bad version, bad sizes
    const std::vector<unsigned char> good = slurp(path);
    check(good.size() > 256, "file size");
    struct Damage {
        const char *what;
        std::vector<unsigned char> bytes;
    };
    std::vector<Damage> bad;
    bad.push_back({"truncated", std::vector<unsigned char>(good.begin(), good.end() - 100)});
    bad.push_back({"flipped bit", good});
    bad.back().bytes[good.size() / 2] ^= 0x10;
    bad.push_back({"bad magic", good});
    bad.back().bytes[0] = 'X';
    bad.push_back({"bad version", good});
    bad.back().bytes[8] = 9;
    bad.push_back({"bad sizes", good});
    const int32_t neg = -5;
    std::memcpy(bad.back().bytes.data() + 16, &neg, 4);
    bad.push_back({"header only", std::vector<unsigned char>(good.begin(), good.begin() + 128)});
    bad.push_back({"short header", std::vector<unsigned char>(good.begin(), good.begin() + 40)});
    for (const Damage &d : bad) {
        const std::string dp = dir + "/selftest_damaged.gpp";
        spit(dp, d.bytes);
        gpp_file_header hh{};
        const gpp_status si = gpp_file_info(dp.c_str(), &hh);
        const gpp_status sr = si ? si
                                 : gpp_file_read(dp.c_str(), out.idx.data(), out.wx.data(), out.vc.data(),
                                                 out.aqsm.data(), out.aqsn.data(), out.ieps.data(), out.wt.data(), nullptr);
        check(sr == GPP_ERR_IO, d.what);
        std::remove(dp.c_str());
    }
    check(gpp_file_info((dir + "/does_not_exist.gpp").c_str(), &h) == GPP_ERR_IO, "missing file");
    // a negative index is refused by the writer
    in.idx[3] = -1;
    check(gpp_file_write(path.c_str(), &p, in.idx.data(), in.wx.data(), 1, in.vc.data(), in.aqsm.data(),
                         in.aqsn.data(), in.ieps.data(), in.wt.data(), nullptr) == GPP_ERR_BAD_INDEX,
          "negative index refused by the writer");
    std::remove(path.c_str());
}

}  // namespace

int main(int argc, char **argv)
{
    const std::string dir = argc > 1 ? argv[1] : ".";
    check(gpp_abi_version() == GPP_ABI_VERSION, "abi version");
    test_fills();
    test_files(dir);
    if (g_failed) {
        std::fprintf(stderr, "gpp_host_selftest: %d checks failed\n", g_failed);
        return 1;
    }
    std::printf("gpp_host_selftest: all checks passed\n");
    return 0;
}
