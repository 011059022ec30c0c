// gpp_io.hip -- problem files (.gpp): host-side writer, header reader,
// streaming payload reader with checksum, and the host-memory reader of the
// C ABI.  The device upload that streams a file through pinned staging
// buffers (gpp_upload_file) is in gpp_api.hip.  Format: gpp_io.hpp.
#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "gpp_io.hpp"

namespace gpp {
namespace io {

namespace {

uint64_t align64(uint64_t x) { return (x + 63) & ~(uint64_t)63; }

void put_u32(unsigned char *b, uint32_t v) { memcpy(b, &v, 4); }
void put_u64(unsigned char *b, uint64_t v) { memcpy(b, &v, 8); }
uint32_t get_u32(const unsigned char *b) { uint32_t v; memcpy(&v, b, 4); return v; }
uint64_t get_u64(const unsigned char *b) { uint64_t v; memcpy(&v, b, 8); return v; }

void encode_header(const Header &h, unsigned char *b)
{
    memset(b, 0, kHeaderBytes);
    memcpy(b, kMagic, 8);
    put_u32(b + 8, kVersion);
    put_u32(b + 12, h.flags);
    const int32_t sz[5] = {h.prob.number_bands, h.prob.nvband, h.prob.ncouls, h.prob.ngpown, h.prob.nfreq};
    memcpy(b + 16, sz, sizeof sz);
    memcpy(b + 40, &h.prob.occ, 8);
    put_u64(b + 48, h.payload);
    put_u64(b + 56, h.checksum);
}

int seek_to(FILE *f, uint64_t off) { return fseeko(f, (off_t)off, SEEK_SET); }

}  // namespace

Layout layout_of(const gpp_problem &p, uint32_t flags)
{
    const uint64_t nb = p.number_bands, nc = p.ncouls, ng = p.ngpown, nf = p.nfreq;
    Layout L;
    memset(&L, 0, sizeof L);
    const uint64_t bytes[kSections] = {
        4 * ng,
        8 * nf * ((flags & 1u) ? nb : 1),
        8 * nc,
        16 * nb * nc,
        16 * nb * nc,
        16 * ng * nc,
        16 * ng * nc,
        (flags & 2u) ? 16 * nf : 0,
        (flags & 2u) ? 16 * nf : 0,
        (flags & 2u) ? 16 * nb : 0,
        (flags & 2u) ? (uint64_t)16 : 0,
    };
    uint64_t off = 0;
    for (int s = 0; s < kSections; ++s) {
        L.bytes[s] = bytes[s];
        if (!bytes[s]) continue;
        L.off[s] = off;
        off = align64(off + bytes[s]);
    }
    L.payload = off;
    return L;
}

gpp_status read_header(void *file, const char *path, Header *h)
{
    FILE *f = static_cast<FILE *>(file);
    unsigned char b[kHeaderBytes];
    if (seek_to(f, 0) != 0 || fread(b, 1, kHeaderBytes, f) != kHeaderBytes)
        return api_error(GPP_ERR_IO, "%s: cannot read the %zu-byte header", path, kHeaderBytes);
    if (memcmp(b, kMagic, 8) != 0) return api_error(GPP_ERR_IO, "%s: not a GPP problem file (bad magic)", path);
    if (get_u32(b + 8) != kVersion)
        return api_error(GPP_ERR_IO, "%s: file version %u, this library reads %u", path, get_u32(b + 8), kVersion);
    h->flags = get_u32(b + 12);
    if (h->flags & ~3u) return api_error(GPP_ERR_IO, "%s: unknown flags 0x%x", path, h->flags);
    int32_t sz[5];
    memcpy(sz, b + 16, sizeof sz);
    h->prob.number_bands = sz[0];
    h->prob.nvband = sz[1];
    h->prob.ncouls = sz[2];
    h->prob.ngpown = sz[3];
    h->prob.nfreq = sz[4];
    memcpy(&h->prob.occ, b + 40, 8);
    h->payload = get_u64(b + 48);
    h->checksum = get_u64(b + 56);
    gpp_status st = api_check_problem(&h->prob);
    if (st) return api_error(GPP_ERR_IO, "%s: header holds invalid sizes (%d %d %d %d %d)", path, sz[0], sz[1], sz[2],
                             sz[3], sz[4]);
    if (h->payload != layout_of(h->prob, h->flags).payload)
        return api_error(GPP_ERR_IO, "%s: payload size %llu does not match the header's sizes", path,
                         (unsigned long long)h->payload);
    return GPP_OK;
}

gpp_status stream_payload(void *file, const char *path, const Header &h, void *const *bufs, int nbufs, size_t cap,
                          Sink &sink)
{
    FILE *f = static_cast<FILE *>(file);
    const Layout L = layout_of(h.prob, h.flags);
    if (seek_to(f, kHeaderBytes) != 0) return api_error(GPP_ERR_IO, "%s: seek failed", path);
    Hash hash;
    uint64_t pos = 0;
    gpp_status st;
    for (int chunk = 0; pos < L.payload; ++chunk) {
        const size_t n = (size_t)std::min<uint64_t>(cap, L.payload - pos);
        const int b = chunk % nbufs;
        if ((st = sink.before_fill(b))) return st;
        unsigned char *buf = static_cast<unsigned char *>(bufs[b]);
        if (fread(buf, 1, n, f) != n)
            return api_error(GPP_ERR_IO, "%s: truncated (payload ends before byte %llu of %llu)", path,
                             (unsigned long long)(pos + n), (unsigned long long)L.payload);
        hash.update(buf, n);
        for (int s = 0; s < kSections; ++s) {
            if (!L.bytes[s]) continue;
            const uint64_t lo = std::max(pos, L.off[s]), hi = std::min<uint64_t>(pos + n, L.off[s] + L.bytes[s]);
            if (lo < hi && (st = sink.piece(s, lo - L.off[s], buf + (lo - pos), hi - lo, b))) return st;
        }
        if ((st = sink.after_chunk(b))) return st;
        pos += n;
    }
    if (hash.value() != h.checksum)
        return api_error(GPP_ERR_IO, "%s: checksum mismatch (stored %016llx, computed %016llx)", path,
                         (unsigned long long)h.checksum, (unsigned long long)hash.value());
    return GPP_OK;
}

}  // namespace io
}  // namespace gpp

using namespace gpp::io;

namespace {

struct FileCloser {
    FILE *f;
    ~FileCloser()
    {
        if (f) fclose(f);
    }
};

// Host destinations of every section.
struct HostSink : Sink {
    void *dst[kSections] = {};
    gpp_status piece(int s, uint64_t off, const void *p, uint64_t n, int) override
    {
        if (dst[s]) memcpy(static_cast<char *>(dst[s]) + off, p, n);
        return GPP_OK;
    }
};

}  // namespace

extern "C" {

gpp_status gpp_file_write(const char *path, const gpp_problem *prob, const int32_t *inv_igp_index,
                          const double *wx, int wx_per_band, const double *vcoul, const gpp_complex *aqsmtemp,
                          const gpp_complex *aqsntemp, const gpp_complex *I_eps_array,
                          const gpp_complex *wtilde_array, const gpp_outputs *expected)
{
    if (!path) return gpp::api_error(GPP_ERR_INVALID_ARGUMENT, "path is NULL");
    gpp_status st = gpp::api_check_problem(prob);
    if (st) return st;
    if (!inv_igp_index || !wx || !vcoul || !aqsmtemp || !aqsntemp || !I_eps_array || !wtilde_array)
        return gpp::api_error(GPP_ERR_INVALID_ARGUMENT, "NULL input array");
    if (expected && (!expected->achtemp || !expected->asxtemp))
        return gpp::api_error(GPP_ERR_INVALID_ARGUMENT, "expected outputs need achtemp and asxtemp");
    for (int j = 0; j < prob->ngpown; ++j)
        if (inv_igp_index[j] < 0)
            return gpp::api_error(GPP_ERR_BAD_INDEX, "inv_igp_index[%d] = %d is negative", j, inv_igp_index[j]);

    Header h;
    h.prob = *prob;
    h.flags = (wx_per_band ? 1u : 0u) | (expected ? 2u : 0u);
    const Layout L = layout_of(h.prob, h.flags);
    h.payload = L.payload;
    h.checksum = 0;

    // NaN where the writer has no expected value.
    std::vector<gpp_complex> nan_acht, nan_achs;
    const void *src[kSections] = {inv_igp_index, wx, vcoul, aqsmtemp, aqsntemp, I_eps_array, wtilde_array};
    if (expected) {
        const gpp_complex nanc = {NAN, NAN};
        src[kAchtemp] = expected->achtemp;
        src[kAsxtemp] = expected->asxtemp;
        if (!expected->acht_n1_loc) nan_acht.assign(prob->number_bands, nanc);
        if (!expected->achstemp) nan_achs.assign(1, nanc);
        src[kAcht] = expected->acht_n1_loc ? (const void *)expected->acht_n1_loc : nan_acht.data();
        src[kAchs] = expected->achstemp ? (const void *)expected->achstemp : nan_achs.data();
    }

    FILE *f = fopen(path, "wb");
    if (!f) return gpp::api_error(GPP_ERR_IO, "%s: cannot open for writing: %s", path, strerror(errno));
    FileCloser closer{f};
    unsigned char hb[kHeaderBytes];
    encode_header(h, hb);
    if (fwrite(hb, 1, kHeaderBytes, f) != kHeaderBytes) return gpp::api_error(GPP_ERR_IO, "%s: write failed", path);
    // Sections go out through one staging block so that the checksum sees
    // whole 64-bit words (padding included).
    const size_t cap = (size_t)16 << 20;
    std::vector<unsigned char> stage(cap);
    Hash hash;
    uint64_t pos = 0;
    size_t fill = 0;
    auto flush = [&]() -> bool {
        hash.update(stage.data(), fill);
        const bool ok = fwrite(stage.data(), 1, fill, f) == fill;
        pos += fill;
        fill = 0;
        return ok;
    };
    for (int s = 0; s < kSections; ++s) {
        if (!L.bytes[s]) continue;
        // zero padding up to the section start
        while (pos + fill < L.off[s]) {
            const size_t n = (size_t)std::min<uint64_t>(L.off[s] - (pos + fill), cap - fill);
            memset(stage.data() + fill, 0, n);
            fill += n;
            if (fill == cap && !flush()) return gpp::api_error(GPP_ERR_IO, "%s: write failed", path);
        }
        const unsigned char *p = static_cast<const unsigned char *>(src[s]);
        uint64_t left = L.bytes[s];
        while (left) {
            const size_t n = (size_t)std::min<uint64_t>(left, cap - fill);
            memcpy(stage.data() + fill, p, n);
            fill += n;
            p += n;
            left -= n;
            if (fill == cap && !flush()) return gpp::api_error(GPP_ERR_IO, "%s: write failed", path);
        }
    }
    if (pos + fill < L.payload) {
        const size_t n = (size_t)(L.payload - (pos + fill));
        memset(stage.data() + fill, 0, n);
        fill += n;
    }
    if (fill && !flush()) return gpp::api_error(GPP_ERR_IO, "%s: write failed", path);
    h.checksum = hash.value();
    encode_header(h, hb);
    if (fseeko(f, 0, SEEK_SET) != 0 || fwrite(hb, 1, kHeaderBytes, f) != kHeaderBytes)
        return gpp::api_error(GPP_ERR_IO, "%s: header rewrite failed", path);
    closer.f = nullptr;
    if (fclose(f) != 0) return gpp::api_error(GPP_ERR_IO, "%s: close failed: %s", path, strerror(errno));
    return GPP_OK;
}

gpp_status gpp_file_info(const char *path, gpp_file_header *info)
{
    if (!path || !info) return gpp::api_error(GPP_ERR_INVALID_ARGUMENT, "NULL argument");
    FILE *f = fopen(path, "rb");
    if (!f) return gpp::api_error(GPP_ERR_IO, "%s: cannot open: %s", path, strerror(errno));
    FileCloser closer{f};
    Header h;
    gpp_status st = read_header(f, path, &h);
    if (st) return st;
    info->problem = h.prob;
    info->flags = h.flags;
    info->payload_bytes = h.payload;
    info->checksum = h.checksum;
    return GPP_OK;
}

gpp_status gpp_file_read(const char *path, int32_t *inv_igp_index, double *wx, double *vcoul,
                         gpp_complex *aqsmtemp, gpp_complex *aqsntemp, gpp_complex *I_eps_array,
                         gpp_complex *wtilde_array, gpp_outputs *expected)
{
    if (!path) return gpp::api_error(GPP_ERR_INVALID_ARGUMENT, "path is NULL");
    if (!inv_igp_index || !wx || !vcoul || !aqsmtemp || !aqsntemp || !I_eps_array || !wtilde_array)
        return gpp::api_error(GPP_ERR_INVALID_ARGUMENT, "NULL output array");
    FILE *f = fopen(path, "rb");
    if (!f) return gpp::api_error(GPP_ERR_IO, "%s: cannot open: %s", path, strerror(errno));
    FileCloser closer{f};
    Header h;
    gpp_status st = read_header(f, path, &h);
    if (st) return st;
    HostSink sink;
    void *dst[kSections] = {inv_igp_index, wx, vcoul, aqsmtemp, aqsntemp, I_eps_array, wtilde_array};
    if (expected && (h.flags & 2u)) {
        dst[kAchtemp] = expected->achtemp;
        dst[kAsxtemp] = expected->asxtemp;
        dst[kAcht] = expected->acht_n1_loc;
        dst[kAchs] = expected->achstemp;
    }
    memcpy(sink.dst, dst, sizeof dst);
    const size_t cap = (size_t)16 << 20;
    std::vector<unsigned char> buf(cap);
    void *bufs[1] = {buf.data()};
    return stream_payload(f, path, h, bufs, 1, cap, sink);
}

}  // extern "C"
