// gpp_io.hpp -- the problem file format (.gpp) shared by the host readers and
// writers of gpp_io.hip and the streaming device upload of gpp_api.hip.
//
// Layout (little-endian):
//   [0, 128)   header
//       0  char     magic[8] = "GPPSIG01"
//       8  uint32   version = 1
//      12  uint32   flags: 1 = wx is [number_bands][nfreq] (band-dependent),
//                          2 = expected outputs follow the inputs
//      16  int32    number_bands, nvband, ncouls, ngpown, nfreq
//      36  int32    0
//      40  double   occ
//      48  uint64   payload bytes
//      56  uint64   payload checksum
//      64  zeros
//   [128, 128 + payload)  sections, each starting at a multiple of 64 bytes
//       (relative to the payload), zero padding between them:
//       inv_igp_index int32[ngpown], wx double[nfreq or nb*nfreq],
//       vcoul double[ncouls], aqsmtemp, aqsntemp complex[nb*ncouls],
//       I_eps_array, wtilde_array complex[ngpown*ncouls]; then, with flag 2,
//       achtemp[nfreq], asxtemp[nfreq], acht_n1_loc[nb], achstemp[1] (complex;
//       NaN where the writer had no value).
// The checksum is FNV-1a over 64-bit words in four interleaved lanes: lane k
// (k = 0..3) starts at h_k = 0xcbf29ce484222325 and takes the little-endian
// payload words w_i with i % 4 == k, h_k = (h_k ^ w_i) * 0x100000001b3
// (mod 2^64), padding included; the checksum is h = 0xcbf29ce484222325 then
// h = (h ^ h_k) * 0x100000001b3 for k = 0..3.  (Four independent chains keep
// the multiplier busy; one chain is latency-bound at about 1 ns per word.)
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "../../include/gpp.h"

namespace gpp {

// Error reporting of the C ABI (thread-local message of gpp_last_error).
gpp_status api_error(gpp_status st, const char *fmt, ...);
gpp_status api_check_problem(const gpp_problem *p);

namespace io {

constexpr char kMagic[8] = {'G', 'P', 'P', 'S', 'I', 'G', '0', '1'};
constexpr uint32_t kVersion = 1;
constexpr size_t kHeaderBytes = 128;
constexpr uint64_t kHashInit = 0xcbf29ce484222325ull;
constexpr uint64_t kHashPrime = 0x100000001b3ull;

enum Section { kIdx, kWx, kVcoul, kAqsm, kAqsn, kIeps, kWtilde, kAchtemp, kAsxtemp, kAcht, kAchs, kSections };

struct Header {
    gpp_problem prob;
    uint32_t flags;
    uint64_t payload;
    uint64_t checksum;
};

struct Layout {
    uint64_t off[kSections];    // offset in the payload (0 for an absent section)
    uint64_t bytes[kSections];  // 0 for an absent section
    uint64_t payload;
};

Layout layout_of(const gpp_problem &p, uint32_t flags);

// Four-lane payload hash.  update() takes whole 32-byte groups (every piece
// the readers and writers pass is a multiple of 64 bytes), so lane k always
// sees the words with index % 4 == k.
struct Hash {
    uint64_t lane[4] = {kHashInit, kHashInit, kHashInit, kHashInit};
    void update(const void *p, size_t n)  // n % 32 == 0
    {
        const unsigned char *b = static_cast<const unsigned char *>(p);
        uint64_t h0 = lane[0], h1 = lane[1], h2 = lane[2], h3 = lane[3];
        for (size_t i = 0; i < n; i += 32) {
            uint64_t w[4];
            __builtin_memcpy(w, b + i, 32);
            h0 = (h0 ^ w[0]) * kHashPrime;
            h1 = (h1 ^ w[1]) * kHashPrime;
            h2 = (h2 ^ w[2]) * kHashPrime;
            h3 = (h3 ^ w[3]) * kHashPrime;
        }
        lane[0] = h0, lane[1] = h1, lane[2] = h2, lane[3] = h3;
    }
    uint64_t value() const
    {
        uint64_t h = kHashInit;
        for (int k = 0; k < 4; ++k) h = (h ^ lane[k]) * kHashPrime;
        return h;
    }
};

// Reads and validates the header of an open file (positioned anywhere).
gpp_status read_header(void *file, const char *path, Header *h);

// Streams the payload of an open file through buf (cap bytes, a multiple of
// 64): for every piece of a section read, calls
// sink(section, offset_in_section, ptr, bytes); then checks the checksum.
// before_fill(i) is called before chunk i is read into buffer `bufs[i % nbufs]`
// so that the caller can wait for its earlier use.
struct Sink {
    virtual ~Sink() = default;
    virtual gpp_status before_fill(int buf) { (void)buf; return GPP_OK; }
    virtual gpp_status piece(int section, uint64_t off, const void *ptr, uint64_t bytes, int buf) = 0;
    virtual gpp_status after_chunk(int buf) { (void)buf; return GPP_OK; }
};
gpp_status stream_payload(void *file, const char *path, const Header &h, void *const *bufs, int nbufs, size_t cap,
                          Sink &sink);

}  // namespace io
}  // namespace gpp
