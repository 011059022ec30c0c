// gpp_api.hip -- the extern "C" boundary of libgpp_hip.so (include/gpp.h).
//
// Host-side only: argument checking, device buffers, HIP streams/events and
// dispatch to the kernel launchers of gpp_direct.hip / gpp_factored.hip.
// Nothing here prints, exits or falls back to a CPU path: every failure is a
// gpp_status plus a thread-local message.
#include <algorithm>
#include <map>
#include <mutex>
#include <new>
#include <tuple>
#include <errno.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <vector>

#include "../../include/gpp.h"
#include "gpp_internal.hpp"
#include "gpp_io.hpp"
#include "gpp_rng.hpp"

using gpp::cplx;

static_assert(sizeof(gpp_complex) == sizeof(cplx), "gpp_complex must be 16 bytes");

namespace {

thread_local char g_err[512] = "";

gpp_status fail(gpp_status st, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return st;
}

gpp_status fail_hip(hipError_t e, const char *what)
{
    gpp_status st = e == hipErrorOutOfMemory ? GPP_ERR_OUT_OF_MEMORY
                  : (e == hipErrorNoDevice || e == hipErrorInvalidDevice) ? GPP_ERR_NO_DEVICE
                                                                          : GPP_ERR_LAUNCH;
    return fail(st, "%s: %s", what, hipGetErrorString(e));
}

#define HIP_OR_FAIL(call, what)                         \
    do {                                                \
        hipError_t e_ = (call);                         \
        if (e_ != hipSuccess) return fail_hip(e_, what); \
    } while (0)

// A device buffer that only grows.
struct DevBuf {
    void *ptr = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes)
    {
        if (bytes <= cap && ptr) return hipSuccess;
        if (ptr) {
            (void)hipFree(ptr);
            ptr = nullptr;
            cap = 0;
        }
        hipError_t e = hipMalloc(&ptr, bytes ? bytes : 256);
        if (e == hipSuccess) cap = bytes;
        else ptr = nullptr;
        return e;
    }
    void release()
    {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
    }
};

gpp_status check_problem(const gpp_problem *p)
{
    if (!p) return fail(GPP_ERR_INVALID_ARGUMENT, "problem is NULL");
    if (p->number_bands < 1 || p->nvband < 0 || p->nvband > p->number_bands || p->ncouls < 1 ||
        p->ngpown < 1 || p->nfreq < 1 || p->nfreq > 4096)
        return fail(GPP_ERR_INVALID_ARGUMENT,
                    "invalid sizes: number_bands=%d nvband=%d ncouls=%d ngpown=%d nfreq=%d "
                    "(need number_bands>=1, 0<=nvband<=number_bands, ncouls>=1, ngpown>=1, 1<=nfreq<=4096)",
                    p->number_bands, p->nvband, p->ncouls, p->ngpown, p->nfreq);
    if ((uint64_t)p->number_bands * (uint64_t)p->ncouls > (1ull << 40) ||
        (uint64_t)p->ngpown * (uint64_t)p->ncouls > (1ull << 40))
        return fail(GPP_ERR_INVALID_ARGUMENT, "problem too large");
    return GPP_OK;
}

bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

}  // namespace

namespace gpp {

gpp_status api_error(gpp_status st, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return st;
}

gpp_status api_check_problem(const gpp_problem *p) { return check_problem(p); }

}  // namespace gpp

// ---------------------------------------------------------------------------
namespace gpp {

hipError_t Workspace::reserve(size_t need)
{
    if (ptr && need <= bytes) return hipSuccess;
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
    hipError_t e = hipMalloc(&ptr, need ? need : 256);
    if (e == hipSuccess) bytes = need;
    else ptr = nullptr;
    return e;
}

void Workspace::release()
{
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

hipError_t StageTimer::init()
{
    for (int k = 0; k < kRing; ++k)
        for (int i = 0; i <= kMax; ++i) {
            hipError_t e = hipEventCreate(&slot[k].ev[i]);
            if (e != hipSuccess) return e;
        }
    ready = true;
    return hipSuccess;
}

void StageTimer::destroy()
{
    for (int k = 0; k < kRing; ++k)
        for (int i = 0; i <= kMax; ++i)
            if (slot[k].ev[i]) (void)hipEventDestroy(slot[k].ev[i]);
    ready = false;
}

void StageTimer::begin(hipStream_t s)
{
    if (filled > 0) cur = (cur + 1) % kRing;
    if (filled < kRing) ++filled;
    slot[cur].n = 0;
    if (ready) (void)hipEventRecord(slot[cur].ev[0], s);
}

void StageTimer::mark(hipStream_t s, const char *nm)
{
    Slot &sl = slot[cur];
    if (!ready || sl.n >= kMax) return;
    sl.name[sl.n] = nm;
    (void)hipEventRecord(sl.ev[sl.n + 1], s);
    ++sl.n;
}

}  // namespace gpp

// ---------------------------------------------------------------------------
struct gpp_ctx {
    int device = 0;
    hipStream_t stream = nullptr;      // where the context enqueues: own_stream or the caller's (gpp_ctx_set_stream)
    hipStream_t own_stream = nullptr;  // created with the context
    int col_j0 = 0, col_j1 = 0;        // gpp_ctx_set_columns: evaluate G' columns [col_j0, col_j1) (0, 0: all)
    gpp_problem prob{};
    bool have_inputs = false;
    gpp::DevProblem dp;
    DevBuf idx, wx, vcoul, aqsm, aqsn, ieps, wtilde;  // owned inputs
    DevBuf results, flag;
    gpp::Workspace work;
    gpp::StageTimer timer;
    bool timed = false;
    int ozaki_slices = 6;
    int ozaki_layout = 4;
    int ozaki_group = 0;
    int gram_tile = 0;
    int red_chunk = 8192;
    int red_fc = 0;  // 0: 16 from 16 frequencies up, else 8
    int red_fpb = 0;
    int fuse_reduce = 0;
    int num_cus = 256;
    int max_buckets = 32;
    // band-dependent frequencies (gpp_set_band_frequencies)
    bool per_band = false;
    DevBuf wxb;                     // [nb][nfreq]
    gpp::BucketPlan bplan;          // bands grouped by identical rows
    gpp_form last_form = GPP_FORM_DIRECT;
    bool have_last_form = false;  // an evaluation was enqueued on the current inputs
    bool last_acht = false;       // ... and it computed acht_n1_loc
    // autotune results: (form, nb, nv, ncouls, ngpown, nfreq) -> (tile, chunk, fc)
    std::map<std::tuple<int, int, int, int, int, int>, std::tuple<int, int, int>> tuned;
};

namespace {

gpp_status set_device(const gpp_ctx *ctx)
{
    HIP_OR_FAIL(hipSetDevice(ctx->device), "hipSetDevice");
    return GPP_OK;
}

void fill_dp_sizes(gpp::DevProblem &dp, const gpp_problem *p)
{
    dp.nb = p->number_bands;
    dp.nv = p->nvband;
    dp.ncouls = p->ncouls;
    dp.ngpown = p->ngpown;
    dp.nfreq = p->nfreq;
    dp.occ = p->occ;
}

gpp_status ensure_owned(gpp_ctx *ctx, const gpp_problem *p)
{
    const size_t nb = p->number_bands, nc = p->ncouls, ng = p->ngpown, nf = p->nfreq;
    HIP_OR_FAIL(ctx->idx.ensure(sizeof(int32_t) * ng), "hipMalloc inv_igp_index");
    HIP_OR_FAIL(ctx->wx.ensure(sizeof(double) * nf), "hipMalloc wx_array");
    HIP_OR_FAIL(ctx->vcoul.ensure(sizeof(double) * nc), "hipMalloc vcoul");
    HIP_OR_FAIL(ctx->aqsm.ensure(sizeof(cplx) * nb * nc), "hipMalloc aqsmtemp");
    HIP_OR_FAIL(ctx->aqsn.ensure(sizeof(cplx) * nb * nc), "hipMalloc aqsntemp");
    HIP_OR_FAIL(ctx->ieps.ensure(sizeof(cplx) * ng * nc), "hipMalloc I_eps_array");
    HIP_OR_FAIL(ctx->wtilde.ensure(sizeof(cplx) * ng * nc), "hipMalloc wtilde_array");
    gpp::DevProblem &dp = ctx->dp;
    fill_dp_sizes(dp, p);
    dp.idx = static_cast<const int32_t *>(ctx->idx.ptr);
    dp.wx = static_cast<const double *>(ctx->wx.ptr);
    dp.vcoul = static_cast<const double *>(ctx->vcoul.ptr);
    dp.aqsm = static_cast<const cplx *>(ctx->aqsm.ptr);
    dp.aqsn = static_cast<const cplx *>(ctx->aqsn.ptr);
    dp.ieps = static_cast<const cplx *>(ctx->ieps.ptr);
    dp.wtilde = static_cast<const cplx *>(ctx->wtilde.ptr);
    return GPP_OK;
}

gpp::GramKind gram_kind(const gpp_ctx *ctx, gpp_form form)
{
    gpp::GramKind k;
    if (form == GPP_FORM_FACTORED_OZAKI || form == GPP_FORM_FACTORED_OZAKI_I8) k.ozaki_slices = ctx->ozaki_slices;
    k.int8 = form == GPP_FORM_FACTORED_OZAKI_I8;
    k.layout = ctx->ozaki_layout;
    k.group = ctx->ozaki_group;
    k.tile = ctx->gram_tile;
    k.red_chunk = ctx->red_chunk;
    k.red_fc = ctx->red_fc;
    k.red_fpb = ctx->red_fpb;
    k.fuse = ctx->fuse_reduce != 0;
    return k;
}

gpp_status check_index_device(gpp_ctx *ctx, const int32_t *d_idx, int ngpown)
{
    HIP_OR_FAIL(ctx->flag.ensure(sizeof(int)), "hipMalloc flag");
    HIP_OR_FAIL(hipMemsetAsync(ctx->flag.ptr, 0, sizeof(int), ctx->stream), "hipMemsetAsync");
    HIP_OR_FAIL(gpp::launch_check_index(d_idx, ngpown, static_cast<int *>(ctx->flag.ptr), ctx->stream),
                "check_index launch");
    int bad = 0;
    HIP_OR_FAIL(hipMemcpyAsync(&bad, ctx->flag.ptr, sizeof(int), hipMemcpyDeviceToHost, ctx->stream),
                "hipMemcpyAsync");
    HIP_OR_FAIL(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    if (bad) return fail(GPP_ERR_BAD_INDEX, "%d negative inv_igp_index entries", bad);
    return GPP_OK;
}

// Whether a device frequency row holds a negative entry (the factored reduce
// picks its kernel by it): one small synchronous copy.
gpp_status wx_negative_device(gpp_ctx *ctx, const double *d_wx, int nfreq, int *neg)
{
    std::vector<double> wx(nfreq);
    HIP_OR_FAIL(hipMemcpyAsync(wx.data(), d_wx, sizeof(double) * nfreq, hipMemcpyDeviceToHost, ctx->stream),
                "D2H wx_array");
    HIP_OR_FAIL(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    *neg = gpp::wx_has_negative(wx.data(), nfreq);
    return GPP_OK;
}

}  // namespace

extern "C" {

const char *gpp_last_error(void) { return g_err; }

int gpp_abi_version(void) { return GPP_ABI_VERSION; }

gpp_status gpp_ctx_create(gpp_ctx **out, int device)
{
    if (!out) return fail(GPP_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) return fail(GPP_ERR_NO_DEVICE, "no HIP device available");
    if (device < 0 || device >= n) return fail(GPP_ERR_NO_DEVICE, "device %d out of range [0,%d)", device, n);
    hipDeviceProp_t prop;
    HIP_OR_FAIL(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(GPP_ERR_NO_DEVICE, "device %d is %s; this library is built for gfx950", device,
                    prop.gcnArchName);
    gpp_ctx *ctx = new (std::nothrow) gpp_ctx();
    if (!ctx) return fail(GPP_ERR_OUT_OF_MEMORY, "host allocation failed");
    ctx->device = device;
    ctx->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking);
    ctx->stream = ctx->own_stream;
    if (e == hipSuccess) e = ctx->timer.init();
    if (e != hipSuccess) {
        gpp_ctx_destroy(ctx);
        return fail_hip(e, "context setup");
    }
    *out = ctx;
    return GPP_OK;
}

gpp_status gpp_ctx_destroy(gpp_ctx *ctx)
{
    if (!ctx) return GPP_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->idx.release(); ctx->wx.release(); ctx->vcoul.release();
    ctx->aqsm.release(); ctx->aqsn.release(); ctx->ieps.release(); ctx->wtilde.release();
    ctx->results.release(); ctx->flag.release(); ctx->wxb.release();
    ctx->work.release();
    ctx->timer.destroy();
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return GPP_OK;
}

gpp_status gpp_upload(gpp_ctx *ctx, const gpp_problem *prob, const int32_t *inv_igp_index,
                      const double *wx_array, const double *vcoul, const gpp_complex *aqsmtemp,
                      const gpp_complex *aqsntemp, const gpp_complex *I_eps_array,
                      const gpp_complex *wtilde_array)
{
    if (!ctx) return fail(GPP_ERR_INVALID_ARGUMENT, "ctx is NULL");
    gpp_status st = check_problem(prob);
    if (st) return st;
    if (!inv_igp_index || !wx_array || !vcoul || !aqsmtemp || !aqsntemp || !I_eps_array || !wtilde_array)
        return fail(GPP_ERR_INVALID_ARGUMENT, "NULL input array");
    for (int j = 0; j < prob->ngpown; ++j)
        if (inv_igp_index[j] < 0)
            return fail(GPP_ERR_BAD_INDEX, "inv_igp_index[%d] = %d is negative", j, inv_igp_index[j]);
    if ((st = set_device(ctx))) return st;
    ctx->have_inputs = false;
    if ((st = ensure_owned(ctx, prob))) return st;
    const size_t nb = prob->number_bands, nc = prob->ncouls, ng = prob->ngpown, nf = prob->nfreq;
    hipStream_t s = ctx->stream;
    HIP_OR_FAIL(hipMemcpyAsync(ctx->idx.ptr, inv_igp_index, sizeof(int32_t) * ng, hipMemcpyHostToDevice, s), "H2D inv_igp_index");
    HIP_OR_FAIL(hipMemcpyAsync(ctx->wx.ptr, wx_array, sizeof(double) * nf, hipMemcpyHostToDevice, s), "H2D wx_array");
    HIP_OR_FAIL(hipMemcpyAsync(ctx->vcoul.ptr, vcoul, sizeof(double) * nc, hipMemcpyHostToDevice, s), "H2D vcoul");
    HIP_OR_FAIL(hipMemcpyAsync(ctx->aqsm.ptr, aqsmtemp, sizeof(cplx) * nb * nc, hipMemcpyHostToDevice, s), "H2D aqsmtemp");
    HIP_OR_FAIL(hipMemcpyAsync(ctx->aqsn.ptr, aqsntemp, sizeof(cplx) * nb * nc, hipMemcpyHostToDevice, s), "H2D aqsntemp");
    HIP_OR_FAIL(hipMemcpyAsync(ctx->ieps.ptr, I_eps_array, sizeof(cplx) * ng * nc, hipMemcpyHostToDevice, s), "H2D I_eps_array");
    HIP_OR_FAIL(hipMemcpyAsync(ctx->wtilde.ptr, wtilde_array, sizeof(cplx) * ng * nc, hipMemcpyHostToDevice, s), "H2D wtilde_array");
    HIP_OR_FAIL(hipStreamSynchronize(s), "upload synchronize");
    ctx->dp.wx_neg = gpp::wx_has_negative(wx_array, (int)nf);
    ctx->prob = *prob;
    ctx->have_inputs = true;
    ctx->per_band = false;
    ctx->have_last_form = false;
    ctx->col_j0 = ctx->col_j1 = 0;
    return GPP_OK;
}

namespace {

// Streams a problem file's sections into the context's device buffers
// through pinned staging buffers (copy of buffer b overlaps the read of the
// other); the small sections are kept on the host.
struct DeviceSink : gpp::io::Sink {
    hipStream_t s = nullptr;
    hipEvent_t ev[2] = {};
    bool used[2] = {false, false};
    void *dev[gpp::io::kSections] = {};
    std::vector<unsigned char> host[gpp::io::kSections];
    gpp_status before_fill(int b) override
    {
        if (used[b]) HIP_OR_FAIL(hipEventSynchronize(ev[b]), "staging wait");
        return GPP_OK;
    }
    gpp_status piece(int sec, uint64_t off, const void *p, uint64_t n, int) override
    {
        if (dev[sec])
            HIP_OR_FAIL(hipMemcpyAsync(static_cast<char *>(dev[sec]) + off, p, n, hipMemcpyHostToDevice, s),
                        "H2D file section");
        else
            memcpy(host[sec].data() + off, p, n);
        return GPP_OK;
    }
    gpp_status after_chunk(int b) override
    {
        HIP_OR_FAIL(hipEventRecord(ev[b], s), "staging event");
        used[b] = true;
        return GPP_OK;
    }
};

struct PinnedStage {
    void *buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    ~PinnedStage()
    {
        for (int i = 0; i < 2; ++i) {
            if (buf[i]) (void)hipHostFree(buf[i]);
            if (ev[i]) (void)hipEventDestroy(ev[i]);
        }
    }
};

}  // namespace

gpp_status gpp_upload_file(gpp_ctx *ctx, const char *path, gpp_outputs *expected)
{
    using namespace gpp::io;
    if (!ctx || !path) return fail(GPP_ERR_INVALID_ARGUMENT, "NULL argument");
    FILE *f = fopen(path, "rb");
    if (!f) return fail(GPP_ERR_IO, "%s: cannot open: %s", path, strerror(errno));
    struct Closer {
        FILE *f;
        ~Closer() { fclose(f); }
    } closer{f};
    Header h;
    gpp_status st = read_header(f, path, &h);
    if (st) return st;
    if ((st = set_device(ctx))) return st;
    ctx->have_inputs = false;
    if ((st = ensure_owned(ctx, &h.prob))) return st;
    const gpp_problem &p = h.prob;
    const Layout L = layout_of(p, h.flags);

    PinnedStage stage;
    const size_t cap = (size_t)32 << 20;
    for (int i = 0; i < 2; ++i) {
        HIP_OR_FAIL(hipHostMalloc(&stage.buf[i], cap, hipHostMallocDefault), "pinned staging allocation");
        HIP_OR_FAIL(hipEventCreateWithFlags(&stage.ev[i], hipEventDisableTiming), "staging event");
    }
    DeviceSink sink;
    sink.s = ctx->stream;
    sink.ev[0] = stage.ev[0];
    sink.ev[1] = stage.ev[1];
    sink.dev[kVcoul] = ctx->vcoul.ptr;
    sink.dev[kAqsm] = ctx->aqsm.ptr;
    sink.dev[kAqsn] = ctx->aqsn.ptr;
    sink.dev[kIeps] = ctx->ieps.ptr;
    sink.dev[kWtilde] = ctx->wtilde.ptr;
    for (int sec : {(int)kIdx, (int)kWx, (int)kAchtemp, (int)kAsxtemp, (int)kAcht, (int)kAchs})
        sink.host[sec].resize(L.bytes[sec]);
    st = stream_payload(f, path, h, stage.buf, 2, cap, sink);
    HIP_OR_FAIL(hipStreamSynchronize(ctx->stream), "file upload synchronize");
    if (st) return st;

    const int32_t *idx = reinterpret_cast<const int32_t *>(sink.host[kIdx].data());
    for (int j = 0; j < p.ngpown; ++j)
        if (idx[j] < 0) return fail(GPP_ERR_BAD_INDEX, "%s: inv_igp_index[%d] = %d is negative", path, j, idx[j]);
    const double *wx = reinterpret_cast<const double *>(sink.host[kWx].data());
    HIP_OR_FAIL(hipMemcpyAsync(ctx->idx.ptr, idx, sizeof(int32_t) * p.ngpown, hipMemcpyHostToDevice, ctx->stream),
                "H2D inv_igp_index");
    HIP_OR_FAIL(hipMemcpyAsync(ctx->wx.ptr, wx, sizeof(double) * p.nfreq, hipMemcpyHostToDevice, ctx->stream),
                "H2D wx_array");
    HIP_OR_FAIL(hipStreamSynchronize(ctx->stream), "file upload synchronize");
    ctx->dp.wx_neg = gpp::wx_has_negative(wx, p.nfreq);
    ctx->prob = p;
    ctx->have_inputs = true;
    ctx->per_band = false;
    ctx->have_last_form = false;
    ctx->col_j0 = ctx->col_j1 = 0;
    if ((h.flags & 1u) && (st = gpp_set_band_frequencies(ctx, wx))) return st;
    if (expected && (h.flags & 2u)) {
        if (expected->achtemp) memcpy(expected->achtemp, sink.host[kAchtemp].data(), L.bytes[kAchtemp]);
        if (expected->asxtemp) memcpy(expected->asxtemp, sink.host[kAsxtemp].data(), L.bytes[kAsxtemp]);
        if (expected->acht_n1_loc) memcpy(expected->acht_n1_loc, sink.host[kAcht].data(), L.bytes[kAcht]);
        if (expected->achstemp) memcpy(expected->achstemp, sink.host[kAchs].data(), L.bytes[kAchs]);
    }
    return GPP_OK;
}

gpp_status gpp_bind_device(gpp_ctx *ctx, const gpp_problem *prob, const int32_t *d_inv_igp_index,
                           const double *d_wx_array, const double *d_vcoul,
                           const gpp_complex *d_aqsmtemp, const gpp_complex *d_aqsntemp,
                           const gpp_complex *d_I_eps_array, const gpp_complex *d_wtilde_array)
{
    if (!ctx) return fail(GPP_ERR_INVALID_ARGUMENT, "ctx is NULL");
    gpp_status st = check_problem(prob);
    if (st) return st;
    if (!d_inv_igp_index || !d_wx_array || !d_vcoul || !d_aqsmtemp || !d_aqsntemp || !d_I_eps_array ||
        !d_wtilde_array)
        return fail(GPP_ERR_INVALID_ARGUMENT, "NULL device array");
    if (!aligned16(d_aqsmtemp) || !aligned16(d_aqsntemp) || !aligned16(d_I_eps_array) || !aligned16(d_wtilde_array))
        return fail(GPP_ERR_INVALID_ARGUMENT, "complex device arrays must be 16-byte aligned");
    if ((st = set_device(ctx))) return st;
    ctx->have_inputs = false;
    if ((st = check_index_device(ctx, d_inv_igp_index, prob->ngpown))) return st;
    int neg = 0;
    if ((st = wx_negative_device(ctx, d_wx_array, prob->nfreq, &neg))) return st;
    gpp::DevProblem &dp = ctx->dp;
    fill_dp_sizes(dp, prob);
    dp.wx_neg = neg;
    dp.idx = d_inv_igp_index;
    dp.wx = d_wx_array;
    dp.vcoul = d_vcoul;
    dp.aqsm = reinterpret_cast<const cplx *>(d_aqsmtemp);
    dp.aqsn = reinterpret_cast<const cplx *>(d_aqsntemp);
    dp.ieps = reinterpret_cast<const cplx *>(d_I_eps_array);
    dp.wtilde = reinterpret_cast<const cplx *>(d_wtilde_array);
    ctx->prob = *prob;
    ctx->have_inputs = true;
    ctx->per_band = false;
    ctx->have_last_form = false;
    ctx->col_j0 = ctx->col_j1 = 0;
    return GPP_OK;
}

gpp_status gpp_fill_deterministic(const gpp_problem *prob, int igp_zero_based, int32_t *inv_igp_index,
                                  double *wx_array, double *vcoul, gpp_complex *aqsmtemp,
                                  gpp_complex *aqsntemp, gpp_complex *I_eps_array,
                                  gpp_complex *wtilde_array)
{
    gpp_status st = check_problem(prob);
    if (st) return st;
    const int nb = prob->number_bands, nc = prob->ncouls, ng = prob->ngpown, nf = prob->nfreq;
    // gppComplex.cpp:239-247
    const double e_lk = 10, dw = 1, to1 = 1e-6, e_n1kq = 6.0;
    // expr = (0.5, 0.5); GPPComplex_mult(expr, (double)(i+j)), :293, :296-310
    for (int i = 0; i < nb; i++)
        for (int j = 0; j < nc; j++) {
            const double v = 0.5 * (double)(i + j);
            const size_t o = (size_t)i * nc + j;
            if (aqsntemp) aqsntemp[o] = gpp_complex{v, v};
            if (aqsmtemp) aqsmtemp[o] = gpp_complex{v, v};
        }
    for (int i = 0; i < ng; i++) {
        for (int j = 0; j < nc; j++) {
            const double v = 0.5 * (double)(i + j);
            const size_t o = (size_t)i * nc + j;
            if (I_eps_array) I_eps_array[o] = gpp_complex{v, v};
            if (wtilde_array) wtilde_array[o] = gpp_complex{v, v};
        }
        // :312
        if (inv_igp_index)
            inv_igp_index[i] = (int32_t)(((int64_t)(igp_zero_based ? i : i + 1) * nc) / ng);
    }
    if (vcoul)
        for (int i = 0; i < nc; i++) vcoul[i] = 1.0 * i;  // :319
    if (wx_array)
        for (int iw = 0; iw < nf; ++iw) {
            wx_array[iw] = e_lk - e_n1kq + dw * ((iw + 1) - 2);  // :327
            if (wx_array[iw] < to1) wx_array[iw] = to1;          // :328
        }
    return GPP_OK;
}

gpp_status gpp_fill_seeded(const gpp_problem *prob, uint64_t seed, double wx_center, double dw,
                           int32_t *inv_igp_index, double *wx_array, double *vcoul, gpp_complex *aqsmtemp,
                           gpp_complex *aqsntemp, gpp_complex *I_eps_array, gpp_complex *wtilde_array)
{
    gpp_status st = check_problem(prob);
    if (st) return st;
    const size_t nb = prob->number_bands, nc = prob->ncouls, ng = prob->ngpown;
    if (ng > nc) return fail(GPP_ERR_INVALID_ARGUMENT, "seeded inputs need ngpown <= ncouls");
    auto fill = [&](gpp_complex *dst, size_t n, uint32_t id, double rlo, double rhi, double ilo, double ihi) {
        if (!dst) return;
        for (size_t i = 0; i < n; ++i) {
            dst[i].re = gpp::seeded_uniform(seed, id, 2 * i, rlo, rhi);
            dst[i].im = gpp::seeded_uniform(seed, id, 2 * i + 1, ilo, ihi);
        }
    };
    fill(aqsmtemp, nb * nc, gpp::kArrAqsm, -1.0, 1.0, -1.0, 1.0);
    fill(aqsntemp, nb * nc, gpp::kArrAqsn, -1.0, 1.0, -1.0, 1.0);
    fill(I_eps_array, ng * nc, gpp::kArrIeps, -0.5, 0.5, -0.5, 0.5);
    fill(wtilde_array, ng * nc, gpp::kArrWtilde, 0.5, 5.0, -0.05, 0.05);
    if (vcoul)
        for (size_t i = 0; i < nc; ++i) vcoul[i] = gpp::seeded_uniform(seed, gpp::kArrVcoul, i, 0.1, 10.0);
    if (inv_igp_index) {
        // Selection sampling (Knuth, TAOCP vol. 2, 3.4.2, Algorithm S): each
        // index is kept with probability needed/remaining; output is sorted.
        size_t need = ng, out = 0;
        for (size_t i = 0; i < nc && need > 0; ++i) {
            const double u = gpp::seeded_u01(seed, gpp::kArrIndex, i);
            if (u * (double)(nc - i) < (double)need) {
                inv_igp_index[out++] = (int32_t)i;
                --need;
            }
        }
    }
    if (wx_array)
        for (int iw = 0; iw < prob->nfreq; ++iw) {
            wx_array[iw] = wx_center + dw * (iw - 0.5 * (prob->nfreq - 1));
            if (wx_array[iw] < 1e-6) wx_array[iw] = 1e-6;
        }
    return GPP_OK;
}

gpp_status gpp_generate_seeded(gpp_ctx *ctx, const gpp_problem *prob, uint64_t seed, double wx_center,
                               double dw)
{
    if (!ctx) return fail(GPP_ERR_INVALID_ARGUMENT, "ctx is NULL");
    gpp_status st = check_problem(prob);
    if (st) return st;
    if (prob->ngpown > prob->ncouls) return fail(GPP_ERR_INVALID_ARGUMENT, "seeded inputs need ngpown <= ncouls");
    if ((st = set_device(ctx))) return st;
    ctx->have_inputs = false;
    if ((st = ensure_owned(ctx, prob))) return st;
    std::vector<int32_t> idx(prob->ngpown);
    std::vector<double> wx(prob->nfreq), vc(prob->ncouls);
    st = gpp_fill_seeded(prob, seed, wx_center, dw, idx.data(), wx.data(), vc.data(), nullptr, nullptr,
                         nullptr, nullptr);
    if (st) return st;
    const size_t nb = prob->number_bands, nc = prob->ncouls, ng = prob->ngpown;
    hipStream_t s = ctx->stream;
    HIP_OR_FAIL(hipMemcpyAsync(ctx->idx.ptr, idx.data(), sizeof(int32_t) * ng, hipMemcpyHostToDevice, s), "H2D idx");
    HIP_OR_FAIL(hipMemcpyAsync(ctx->wx.ptr, wx.data(), sizeof(double) * prob->nfreq, hipMemcpyHostToDevice, s), "H2D wx");
    HIP_OR_FAIL(hipMemcpyAsync(ctx->vcoul.ptr, vc.data(), sizeof(double) * nc, hipMemcpyHostToDevice, s), "H2D vcoul");
    HIP_OR_FAIL(gpp::launch_fill_seeded(static_cast<cplx *>(ctx->aqsm.ptr), nb * nc, seed, gpp::kArrAqsm, -1.0, 1.0, -1.0, 1.0, s), "fill aqsm");
    HIP_OR_FAIL(gpp::launch_fill_seeded(static_cast<cplx *>(ctx->aqsn.ptr), nb * nc, seed, gpp::kArrAqsn, -1.0, 1.0, -1.0, 1.0, s), "fill aqsn");
    HIP_OR_FAIL(gpp::launch_fill_seeded(static_cast<cplx *>(ctx->ieps.ptr), ng * nc, seed, gpp::kArrIeps, -0.5, 0.5, -0.5, 0.5, s), "fill I_eps");
    HIP_OR_FAIL(gpp::launch_fill_seeded(static_cast<cplx *>(ctx->wtilde.ptr), ng * nc, seed, gpp::kArrWtilde, 0.5, 5.0, -0.05, 0.05, s), "fill wtilde");
    HIP_OR_FAIL(hipStreamSynchronize(s), "generate synchronize");
    ctx->dp.wx_neg = gpp::wx_has_negative(wx.data(), prob->nfreq);
    ctx->prob = *prob;
    ctx->have_inputs = true;
    ctx->per_band = false;
    ctx->have_last_form = false;
    ctx->col_j0 = ctx->col_j1 = 0;
    return GPP_OK;
}

gpp_status gpp_set_band_frequencies(gpp_ctx *ctx, const double *wx_bands)
{
    if (!ctx) return fail(GPP_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (!ctx->have_inputs) return fail(GPP_ERR_NO_INPUTS, "no inputs: call gpp_upload or gpp_bind_device first");
    if (!wx_bands) {
        ctx->per_band = false;
        return GPP_OK;
    }
    const int nb = ctx->dp.nb, nf = ctx->dp.nfreq;
    for (size_t i = 0; i < (size_t)nb * nf; ++i)
        if (!(wx_bands[i] == wx_bands[i]))
            return fail(GPP_ERR_INVALID_ARGUMENT, "wx_bands[%zu] is NaN", i);
    gpp_status st = set_device(ctx);
    if (st) return st;
    HIP_OR_FAIL(ctx->wxb.ensure(sizeof(double) * nb * nf), "hipMalloc wx_bands");
    HIP_OR_FAIL(hipMemcpyAsync(ctx->wxb.ptr, wx_bands, sizeof(double) * nb * nf, hipMemcpyHostToDevice, ctx->stream),
                "H2D wx_bands");
    HIP_OR_FAIL(hipStreamSynchronize(ctx->stream), "wx_bands synchronize");
    ctx->bplan = gpp::plan_buckets(wx_bands, nb, ctx->dp.nv, nf);
    ctx->per_band = true;
    return GPP_OK;
}

gpp_status gpp_ctx_band_buckets(const gpp_ctx *ctx, int *nbuckets)
{
    if (!ctx || !nbuckets) return fail(GPP_ERR_INVALID_ARGUMENT, "NULL argument");
    *nbuckets = ctx->per_band ? (int)ctx->bplan.row.size() : 1;
    return GPP_OK;
}

gpp_status gpp_ctx_last_form(const gpp_ctx *ctx, gpp_form *form)
{
    if (!ctx || !form) return fail(GPP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!ctx->have_last_form) return fail(GPP_ERR_NO_INPUTS, "no gpp_sigma has run on this context");
    *form = ctx->last_form;
    return GPP_OK;
}

gpp_status gpp_ctx_set_stream(gpp_ctx *ctx, struct ihipStream_t *stream)
{
    if (!ctx) return fail(GPP_ERR_INVALID_ARGUMENT, "ctx is NULL");
    gpp_status st = set_device(ctx);
    if (st) return st;
    hipStream_t next = stream ? stream : ctx->own_stream;
    if (next == ctx->stream) return GPP_OK;
    if (stream) {  // the stream must belong to the context's device
        int dev = -1;
        HIP_OR_FAIL(hipStreamGetDevice(stream, &dev), "hipStreamGetDevice");
        if (dev != ctx->device)
            return fail(GPP_ERR_INVALID_ARGUMENT, "stream is on device %d, the context on device %d", dev,
                        ctx->device);
    }
    // the context's earlier work (results, timing events) finishes first
    HIP_OR_FAIL(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    ctx->stream = next;
    return GPP_OK;
}

gpp_status gpp_ctx_set_columns(gpp_ctx *ctx, int j0, int j1)
{
    if (!ctx) return fail(GPP_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (!ctx->have_inputs) return fail(GPP_ERR_NO_INPUTS, "no inputs: call gpp_upload or gpp_bind_device first");
    if (j0 == 0 && j1 == 0) {
        ctx->col_j0 = ctx->col_j1 = 0;
        return GPP_OK;
    }
    if (j0 < 0 || j1 <= j0 || j1 > ctx->dp.ngpown)
        return fail(GPP_ERR_INVALID_ARGUMENT, "columns [%d, %d) not within [0, %d)", j0, j1, ctx->dp.ngpown);
    ctx->col_j0 = j0;
    ctx->col_j1 = j1;
    ctx->have_last_form = false;
    return GPP_OK;
}

}  // extern "C"

namespace {

// Enqueues one evaluation on the context's stream; the results stay in the
// context's device result buffer.
gpp_status sigma_enqueue(gpp_ctx *ctx, gpp_form form, bool acht)
{
    if (!ctx->have_inputs) return fail(GPP_ERR_NO_INPUTS, "no inputs: call gpp_upload or gpp_bind_device first");
    if (form < GPP_FORM_DIRECT || form > GPP_FORM_FACTORED_OZAKI_I8)
        return fail(GPP_ERR_INVALID_ARGUMENT, "unknown form %d", (int)form);
    gpp_status st = set_device(ctx);
    if (st) return st;
    // Band-dependent frequencies: the direct forms read a row per band; the
    // factored forms factor each group of bands sharing a row (one group:
    // the plain factored form on that row) and fall back to the direct form
    // beyond max_buckets groups.
    gpp::DevProblem dp = ctx->dp;
    if (ctx->col_j1 > ctx->col_j0) {  // one slab of G' columns (gpp_ctx_set_columns)
        const size_t off = (size_t)ctx->col_j0 * dp.ncouls;
        dp.ngpown = ctx->col_j1 - ctx->col_j0;
        dp.idx += ctx->col_j0;
        dp.ieps += off;
        dp.wtilde += off;
    }
    bool bucketed = false;
    if (ctx->per_band) {
        const int nbk = (int)ctx->bplan.row.size();
        const bool factored = form != GPP_FORM_DIRECT && form != GPP_FORM_DIRECT_STRICT;
        if (factored && nbk == 1) {
            dp.wx = static_cast<const double *>(ctx->wxb.ptr);
        } else {
            dp.wx = static_cast<const double *>(ctx->wxb.ptr);
            dp.wx_stride = dp.nfreq;
            if (factored && nbk <= ctx->max_buckets) bucketed = true;
            else if (factored) form = GPP_FORM_DIRECT;
        }
    }
    const bool direct = form == GPP_FORM_DIRECT || form == GPP_FORM_DIRECT_STRICT;
    size_t need = 0;
    gpp::GramKind kind;
    if (direct) {
        need = gpp::direct_workspace_bytes(dp);
    } else {
        kind = gram_kind(ctx, form);
        kind.acht = acht;
        need = bucketed ? gpp::bucketed_workspace_bytes(dp, kind, ctx->bplan) : gpp::factored_workspace_bytes(dp, kind);
    }
    if (need == 0) return fail(GPP_ERR_UNSUPPORTED, "form %d is not available for this problem", (int)form);
    HIP_OR_FAIL(ctx->work.reserve(need), "workspace allocation");
    gpp::DevResults res;
    res.nfreq = dp.nfreq;
    res.nb = dp.nb;
    HIP_OR_FAIL(ctx->results.ensure(sizeof(cplx) * res.count()), "results allocation");
    res.base = static_cast<cplx *>(ctx->results.ptr);
    hipError_t e;
    if (direct)
        e = gpp::run_direct(dp, form == GPP_FORM_DIRECT_STRICT, ctx->work.ptr, res, ctx->stream, &ctx->timer);
    else if (bucketed)
        e = gpp::run_factored_bucketed(dp, kind, ctx->bplan, ctx->work.ptr, res, ctx->stream, &ctx->timer);
    else
        e = gpp::run_factored(dp, kind, ctx->work.ptr, res, ctx->stream, &ctx->timer);
    if (e != hipSuccess) {
        ctx->have_last_form = false;
        return fail_hip(e, "sigma launch");
    }
    ctx->last_form = form;
    ctx->have_last_form = true;
    ctx->last_acht = direct || acht;
    ctx->timed = true;
    return GPP_OK;
}

}  // namespace

extern "C" {

gpp_status gpp_sigma_async(gpp_ctx *ctx, gpp_form form, int with_acht_n1_loc)
{
    if (!ctx) return fail(GPP_ERR_INVALID_ARGUMENT, "ctx is NULL");
    return sigma_enqueue(ctx, form, with_acht_n1_loc != 0);
}

gpp_status gpp_ctx_wait(gpp_ctx *ctx, const gpp_outputs *out)
{
    if (!ctx) return fail(GPP_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (!ctx->have_last_form) return fail(GPP_ERR_NO_INPUTS, "no evaluation enqueued on the current inputs");
    if (out && out->acht_n1_loc && !ctx->last_acht)
        return fail(GPP_ERR_UNSUPPORTED, "the last evaluation did not compute acht_n1_loc (request it when enqueuing)");
    gpp_status st = set_device(ctx);
    if (st) return st;
    const size_t nf = ctx->dp.nfreq, nb = ctx->dp.nb, count = 2 * nf + nb + 1;
    if (!out) {
        HIP_OR_FAIL(hipStreamSynchronize(ctx->stream), "sigma synchronize");
        return GPP_OK;
    }
    std::vector<cplx> host(count);
    HIP_OR_FAIL(hipMemcpyAsync(host.data(), ctx->results.ptr, sizeof(cplx) * count, hipMemcpyDeviceToHost,
                               ctx->stream),
                "D2H results");
    HIP_OR_FAIL(hipStreamSynchronize(ctx->stream), "sigma synchronize");
    if (out->achtemp) memcpy(out->achtemp, host.data(), sizeof(cplx) * nf);
    if (out->asxtemp) memcpy(out->asxtemp, host.data() + nf, sizeof(cplx) * nf);
    if (out->acht_n1_loc) memcpy(out->acht_n1_loc, host.data() + 2 * nf, sizeof(cplx) * nb);
    if (out->achstemp) memcpy(out->achstemp, host.data() + 2 * nf + nb, sizeof(cplx));
    return GPP_OK;
}

gpp_status gpp_sigma(gpp_ctx *ctx, gpp_form form, const gpp_outputs *out)
{
    if (!ctx) return fail(GPP_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (!out || !out->achtemp || !out->asxtemp)
        return fail(GPP_ERR_INVALID_ARGUMENT, "outputs achtemp and asxtemp are required");
    gpp_status st = sigma_enqueue(ctx, form, out->acht_n1_loc != nullptr);
    if (st) return st;
    return gpp_ctx_wait(ctx, out);
}

gpp_status gpp_ctx_set_option(gpp_ctx *ctx, gpp_option option, int value)
{
    if (!ctx) return fail(GPP_ERR_INVALID_ARGUMENT, "ctx is NULL");
    switch (option) {
    case GPP_OPT_OZAKI_SLICES:
        if (value < 3 || value > 6) return fail(GPP_ERR_INVALID_ARGUMENT, "ozaki slices must be 3..6, got %d", value);
        ctx->ozaki_slices = value;
        return GPP_OK;
    case GPP_OPT_OZAKI_LAYOUT:
        if (value < 0 || value > 4) return fail(GPP_ERR_INVALID_ARGUMENT, "ozaki layout must be 0..4, got %d", value);
        ctx->ozaki_layout = value;
        return GPP_OK;
    case GPP_OPT_OZAKI_GROUP:
        if (value < 0 || value > (1 << 16))
            return fail(GPP_ERR_INVALID_ARGUMENT, "ozaki group must be in [0, 65536], got %d", value);
        ctx->ozaki_group = value;
        return GPP_OK;
    case GPP_OPT_GRAM_TILE:
        if (value < 0 || value >= gpp::kGramTiles)
            return fail(GPP_ERR_INVALID_ARGUMENT, "gram tile must be 0..%d, got %d", gpp::kGramTiles - 1, value);
        ctx->gram_tile = value;
        return GPP_OK;
    case GPP_OPT_REDUCE_CHUNK:
        if (value < 256 || value > (1 << 20) || (value & (value - 1)))
            return fail(GPP_ERR_INVALID_ARGUMENT, "reduce chunk must be a power of two in [256, 2^20], got %d", value);
        ctx->red_chunk = value;
        return GPP_OK;
    case GPP_OPT_REDUCE_FREQS:
        if (value != 0 && value != 4 && value != 8 && value != 16)
            return fail(GPP_ERR_INVALID_ARGUMENT, "reduce freqs must be 0, 4, 8 or 16, got %d", value);
        ctx->red_fc = value;
        return GPP_OK;
    case GPP_OPT_REDUCE_FREQS_PER_BLOCK:
        if (value < 0 || value > 4096)
            return fail(GPP_ERR_INVALID_ARGUMENT, "reduce freqs per block must be 0..4096, got %d", value);
        ctx->red_fpb = value;
        return GPP_OK;
    case GPP_OPT_FUSED_REDUCE:
        if (value != 0 && value != 1) return fail(GPP_ERR_INVALID_ARGUMENT, "fused reduce must be 0 or 1, got %d", value);
        ctx->fuse_reduce = value;
        return GPP_OK;
    case GPP_OPT_MAX_BUCKETS:
        if (value < 1 || value > (1 << 20))
            return fail(GPP_ERR_INVALID_ARGUMENT, "max buckets must be in [1, 2^20], got %d", value);
        ctx->max_buckets = value;
        return GPP_OK;
    default: return fail(GPP_ERR_INVALID_ARGUMENT, "unknown option %d", (int)option);
    }
}

gpp_status gpp_ctx_get_option(const gpp_ctx *ctx, gpp_option option, int *value)
{
    if (!ctx || !value) return fail(GPP_ERR_INVALID_ARGUMENT, "NULL argument");
    switch (option) {
    case GPP_OPT_OZAKI_SLICES: *value = ctx->ozaki_slices; return GPP_OK;
    case GPP_OPT_OZAKI_LAYOUT: *value = ctx->ozaki_layout; return GPP_OK;
    case GPP_OPT_OZAKI_GROUP: *value = ctx->ozaki_group; return GPP_OK;
    case GPP_OPT_GRAM_TILE: *value = ctx->gram_tile; return GPP_OK;
    case GPP_OPT_REDUCE_CHUNK: *value = ctx->red_chunk; return GPP_OK;
    case GPP_OPT_REDUCE_FREQS: *value = ctx->red_fc; return GPP_OK;
    case GPP_OPT_MAX_BUCKETS: *value = ctx->max_buckets; return GPP_OK;
    case GPP_OPT_REDUCE_FREQS_PER_BLOCK: *value = ctx->red_fpb; return GPP_OK;
    case GPP_OPT_FUSED_REDUCE: *value = ctx->fuse_reduce; return GPP_OK;
    default: return fail(GPP_ERR_INVALID_ARGUMENT, "unknown option %d", (int)option);
    }
}

gpp_status gpp_ctx_autotune(gpp_ctx *ctx, gpp_form form)
{
    if (!ctx) return fail(GPP_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (form == GPP_FORM_DIRECT || form == GPP_FORM_DIRECT_STRICT) return GPP_OK;
    if (form != GPP_FORM_FACTORED_F64 && form != GPP_FORM_FACTORED_OZAKI && form != GPP_FORM_FACTORED_OZAKI_I8)
        return fail(GPP_ERR_INVALID_ARGUMENT, "unknown form %d", (int)form);
    if (!ctx->have_inputs) return fail(GPP_ERR_NO_INPUTS, "no inputs: call gpp_upload or gpp_bind_device first");
    const gpp::DevProblem &dp = ctx->dp;
    const auto key = std::make_tuple((int)form, dp.nb, dp.nv, dp.ncouls, dp.ngpown, dp.nfreq);
    auto hit = ctx->tuned.find(key);
    if (hit != ctx->tuned.end()) {
        std::tie(ctx->gram_tile, ctx->red_chunk, ctx->red_fc) = hit->second;
        return GPP_OK;
    }
    gpp_status st = set_device(ctx);
    if (st) return st;
    gpp::DevResults res;
    res.nfreq = dp.nfreq;
    res.nb = dp.nb;
    HIP_OR_FAIL(ctx->results.ensure(sizeof(cplx) * res.count()), "results allocation");
    res.base = static_cast<cplx *>(ctx->results.ptr);
    // time one configuration: returns the Gram and reduce stage times (ms)
    auto trial = [&](int tile, int chunk, int fc, double *gram_ms, double *red_ms) -> gpp_status {
        gpp::GramKind k = gram_kind(ctx, form);
        k.tile = tile;
        k.red_chunk = chunk;
        k.red_fc = fc;
        HIP_OR_FAIL(ctx->work.reserve(gpp::factored_workspace_bytes(dp, k)), "workspace allocation");
        double best_g = 1e300, best_r = 1e300;
        for (int rep = 0; rep < 3; ++rep) {
            HIP_OR_FAIL(gpp::run_factored(dp, k, ctx->work.ptr, res, ctx->stream, &ctx->timer), "autotune launch");
            HIP_OR_FAIL(hipStreamSynchronize(ctx->stream), "autotune synchronize");
            if (rep == 0) continue;  // warm-up
            float ms = 0.f;
            double g_ms = 0.0, r_ms = 0.0;
            const gpp::StageTimer::Slot &t = ctx->timer.last();
            for (int i = 0; i < t.n; ++i) {
                HIP_OR_FAIL(hipEventElapsedTime(&ms, t.ev[i], t.ev[i + 1]), "hipEventElapsedTime");
                if (strncmp(t.name[i], "factored_gram", 13) == 0) g_ms = ms;
                if (strcmp(t.name[i], "factored_pole_reduce") == 0 || strcmp(t.name[i], "factored_pole_cols") == 0)
                    r_ms = ms;
            }
            best_g = std::min(best_g, g_ms);
            best_r = std::min(best_r, r_ms);
        }
        *gram_ms = best_g;
        *red_ms = best_r;
        return GPP_OK;
    };
    int best_tile = ctx->gram_tile, best_chunk = ctx->red_chunk, best_fc = ctx->red_fc;
    double g, r, best = 1e300;
    if (form == GPP_FORM_FACTORED_F64)
        for (int tile = 0; tile < gpp::kGramTiles; ++tile) {
            // Gram plus reduce (with GPP_OPT_FUSED_REDUCE the 128x128 tile
            // carries the reduce in its epilogue where it applies)
            if ((st = trial(tile, best_chunk, best_fc, &g, &r))) return st;
            if (g + r < best) best = g + r, best_tile = tile;
        }
    best = 1e300;
    const int chunks[] = {2048, 4096, 8192, 16384};
    for (int chunk : chunks)
        for (int fc : {4, 8, 16}) {
            if (dp.nfreq <= 4 && fc != 4) continue;  // identical launches for nfreq <= 4
            if (dp.nfreq < 16 && fc == 16) continue;  // a 16-wide pass would run idle frequencies
            if ((st = trial(best_tile, chunk, fc, &g, &r))) return st;
            if (r < best) best = r, best_chunk = chunk, best_fc = fc;
        }
    ctx->gram_tile = best_tile;
    ctx->red_chunk = best_chunk;
    ctx->red_fc = best_fc;
    ctx->tuned[key] = std::make_tuple(best_tile, best_chunk, best_fc);
    return GPP_OK;
}

gpp_status gpp_ctx_last_timing(const gpp_ctx *ctx, double *total_ms, double *stage_ms, int max_stages,
                               int *nstages)
{
    if (!ctx) return fail(GPP_ERR_INVALID_ARGUMENT, "ctx is NULL");
    const gpp::StageTimer::Slot &t = ctx->timer.last();
    if (!ctx->timed || t.n == 0) {
        if (total_ms) *total_ms = 0.0;
        if (nstages) *nstages = 0;
        return GPP_OK;
    }
    HIP_OR_FAIL(hipEventSynchronize(t.ev[t.n]), "hipEventSynchronize");
    float ms = 0.f;
    HIP_OR_FAIL(hipEventElapsedTime(&ms, t.ev[0], t.ev[t.n]), "hipEventElapsedTime");
    if (total_ms) *total_ms = ms;
    for (int i = 0; i < t.n && stage_ms && i < max_stages; ++i) {
        HIP_OR_FAIL(hipEventElapsedTime(&ms, t.ev[i], t.ev[i + 1]), "hipEventElapsedTime");
        stage_ms[i] = ms;
    }
    if (nstages) *nstages = t.n;
    return GPP_OK;
}

gpp_status gpp_ctx_timing_summary(const gpp_ctx *ctx, int last_n, int *nevals, double *avg_total_ms,
                                  double *avg_stage_ms, int max_stages, int *nstages)
{
    if (!ctx || last_n < 1) return fail(GPP_ERR_INVALID_ARGUMENT, "ctx is NULL or last_n < 1");
    const gpp::StageTimer &tm = ctx->timer;
    const gpp::StageTimer::Slot &L = tm.last();
    if (nevals) *nevals = 0;
    if (avg_total_ms) *avg_total_ms = 0.0;
    if (nstages) *nstages = 0;
    if (!ctx->timed || L.n == 0) return GPP_OK;
    HIP_OR_FAIL(hipEventSynchronize(L.ev[L.n]), "hipEventSynchronize");
    const int want = std::min(last_n, tm.filled);
    double tot = 0.0, st[gpp::StageTimer::kMax] = {};
    int m = 0;
    for (int k = 0; k < want; ++k) {
        const gpp::StageTimer::Slot &S = tm.back(k);
        bool same = S.n == L.n;
        for (int i = 0; same && i < L.n; ++i) same = S.name[i] && strcmp(S.name[i], L.name[i]) == 0;
        if (!same) break;  // an evaluation of another kind: stop at it
        float ms = 0.f;
        HIP_OR_FAIL(hipEventElapsedTime(&ms, S.ev[0], S.ev[S.n]), "hipEventElapsedTime");
        tot += ms;
        for (int i = 0; i < S.n; ++i) {
            HIP_OR_FAIL(hipEventElapsedTime(&ms, S.ev[i], S.ev[i + 1]), "hipEventElapsedTime");
            st[i] += ms;
        }
        ++m;
    }
    if (nevals) *nevals = m;
    if (avg_total_ms) *avg_total_ms = tot / m;
    for (int i = 0; i < L.n && avg_stage_ms && i < max_stages; ++i) avg_stage_ms[i] = st[i] / m;
    if (nstages) *nstages = L.n;
    return GPP_OK;
}

const char *gpp_ctx_stage_name(const gpp_ctx *ctx, int i)
{
    if (!ctx) return "";
    const gpp::StageTimer::Slot &t = ctx->timer.last();
    if (i < 0 || i >= t.n || !t.name[i]) return "";
    return t.name[i];
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batches.
struct gpp_batch {
    gpp_ctx *ctx = nullptr;
    int n = 0;
    std::vector<gpp::DevProblem> dps;
    std::vector<DevBuf> owned;  // library-owned inputs (7 per instance) for seeded batches
    gpp::GramKind kind;
    bool planned = false;
    gpp::BatchPlan plan;
    DevBuf work;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool use_graph = true;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t evs[3] = {nullptr, nullptr, nullptr};
    double last_ms = 0.0;
    double stage_ms[4] = {0, 0, 0, 0};
    cplx *host_out = nullptr;  // pinned copy of the results table
    size_t host_out_bytes = 0;
};

namespace {

int batch_blocks(const gpp_batch *b) { return 2 * b->ctx->num_cus; }

// The batch's GramKind for `form`: its Gram tile and reduce schedule, with
// the Gram product of the form (FP64 MFMA, or the Ozaki split with the
// context's slice count).
gpp_status batch_kind(const gpp_batch *b, gpp_form form, gpp::GramKind *k)
{
    *k = b->kind;
    k->ozaki_slices = 0;
    k->int8 = false;
    if (form == GPP_FORM_FACTORED_OZAKI || form == GPP_FORM_FACTORED_OZAKI_I8) {
        k->ozaki_slices = b->ctx->ozaki_slices;
        k->int8 = form == GPP_FORM_FACTORED_OZAKI_I8;
    } else if (form != GPP_FORM_FACTORED_F64) {
        return fail(GPP_ERR_UNSUPPORTED, "batches run the factored forms only");
    }
    return GPP_OK;
}

void batch_drop_graph(gpp_batch *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

gpp_status batch_prepare(gpp_batch *b, const gpp::GramKind &k)
{
    batch_drop_graph(b);
    size_t need = 0;
    b->plan = gpp::plan_batch(b->dps, k, &need);
    HIP_OR_FAIL(b->work.ensure(need), "batch workspace allocation");
    HIP_OR_FAIL(gpp::upload_batch(b->plan, b->dps, b->work.ptr, b->ctx->stream), "batch tables upload");
    if (b->plan.out_bytes > b->host_out_bytes) {
        if (b->host_out) (void)hipHostFree(b->host_out);
        b->host_out = nullptr;
        b->host_out_bytes = 0;
        HIP_OR_FAIL(hipHostMalloc(reinterpret_cast<void **>(&b->host_out), b->plan.out_bytes, hipHostMallocDefault),
                    "pinned results allocation");
        b->host_out_bytes = b->plan.out_bytes;
    }
    b->kind = k;
    b->planned = true;
    return GPP_OK;
}

gpp_status batch_capture(gpp_batch *b)
{
    hipStream_t s = b->ctx->stream;
    HIP_OR_FAIL(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    hipError_t e = gpp::launch_batch(b->plan, b->work.ptr, batch_blocks(b), s, 15);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(s, &g);
    if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail_hip(e, "batch capture launch");
    }
    if (e2 != hipSuccess) return fail_hip(e2, "hipStreamEndCapture");
    b->graph = g;
    HIP_OR_FAIL(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0), "hipGraphInstantiate");
    return GPP_OK;
}

gpp_status batch_new(gpp_ctx *ctx, int n, const gpp_problem *probs, gpp_batch **out, gpp_batch **made)
{
    if (!ctx || !out) return fail(GPP_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    if (n < 1 || !probs) return fail(GPP_ERR_INVALID_ARGUMENT, "a batch needs n >= 1 problems");
    for (int i = 0; i < n; ++i) {
        gpp_status st = check_problem(&probs[i]);
        if (st) return st;
    }
    gpp_status st = set_device(ctx);
    if (st) return st;
    gpp_batch *b = new (std::nothrow) gpp_batch();
    if (!b) return fail(GPP_ERR_OUT_OF_MEMORY, "host allocation failed");
    b->ctx = ctx;
    b->n = n;
    b->dps.resize(n);
    for (int i = 0; i < n; ++i) fill_dp_sizes(b->dps[i], &probs[i]);
    b->kind.tile = ctx->gram_tile;
    b->kind.red_chunk = ctx->red_chunk;
    b->kind.red_fc = ctx->red_fc;
    if (hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess ||
        hipEventCreate(&b->evs[0]) != hipSuccess || hipEventCreate(&b->evs[1]) != hipSuccess ||
        hipEventCreate(&b->evs[2]) != hipSuccess) {
        gpp_batch_destroy(b);
        return fail(GPP_ERR_LAUNCH, "hipEventCreate failed");
    }
    *made = b;
    return GPP_OK;
}

}  // namespace

extern "C" {

gpp_status gpp_batch_create(gpp_ctx *ctx, int n, const gpp_problem *probs, const gpp_device_inputs *inputs,
                            gpp_batch **out)
{
    if (!inputs) return fail(GPP_ERR_INVALID_ARGUMENT, "inputs is NULL");
    gpp_batch *b = nullptr;
    gpp_status st = batch_new(ctx, n, probs, out, &b);
    if (st) return st;
    for (int i = 0; i < n; ++i) {
        const gpp_device_inputs &in = inputs[i];
        if (!in.inv_igp_index || !in.wx_array || !in.vcoul || !in.aqsmtemp || !in.aqsntemp || !in.I_eps_array ||
            !in.wtilde_array || !aligned16(in.aqsmtemp) || !aligned16(in.aqsntemp) || !aligned16(in.I_eps_array) ||
            !aligned16(in.wtilde_array)) {
            gpp_batch_destroy(b);
            return fail(GPP_ERR_INVALID_ARGUMENT, "instance %d: NULL or misaligned device array", i);
        }
        int neg = 0;
        if ((st = check_index_device(ctx, in.inv_igp_index, probs[i].ngpown)) ||
            (st = wx_negative_device(ctx, in.wx_array, probs[i].nfreq, &neg))) {
            gpp_batch_destroy(b);
            return st;
        }
        gpp::DevProblem &dp = b->dps[i];
        dp.wx_neg = neg;
        dp.idx = in.inv_igp_index;
        dp.wx = in.wx_array;
        dp.vcoul = in.vcoul;
        dp.aqsm = reinterpret_cast<const cplx *>(in.aqsmtemp);
        dp.aqsn = reinterpret_cast<const cplx *>(in.aqsntemp);
        dp.ieps = reinterpret_cast<const cplx *>(in.I_eps_array);
        dp.wtilde = reinterpret_cast<const cplx *>(in.wtilde_array);
    }
    *out = b;
    return GPP_OK;
}

gpp_status gpp_batch_create_seeded(gpp_ctx *ctx, int n, const gpp_problem *probs, const uint64_t *seeds,
                                   const double *wx_center, const double *dw, gpp_batch **out)
{
    if (!seeds) return fail(GPP_ERR_INVALID_ARGUMENT, "seeds is NULL");
    gpp_batch *b = nullptr;
    gpp_status st = batch_new(ctx, n, probs, out, &b);
    if (st) return st;
    b->owned.resize((size_t)7 * n);
    hipStream_t s = ctx->stream;
    for (int i = 0; i < n; ++i) {
        const gpp_problem &p = probs[i];
        if (p.ngpown > p.ncouls) {
            gpp_batch_destroy(b);
            return fail(GPP_ERR_INVALID_ARGUMENT, "instance %d: seeded inputs need ngpown <= ncouls", i);
        }
        const size_t nb = p.number_bands, nc = p.ncouls, ng = p.ngpown;
        DevBuf *o = &b->owned[(size_t)7 * i];
        const size_t bytes[7] = {sizeof(int32_t) * ng, sizeof(double) * p.nfreq, sizeof(double) * nc,
                                 sizeof(cplx) * nb * nc, sizeof(cplx) * nb * nc, sizeof(cplx) * ng * nc,
                                 sizeof(cplx) * ng * nc};
        for (int q = 0; q < 7; ++q) {
            hipError_t e = o[q].ensure(bytes[q]);
            if (e != hipSuccess) {
                gpp_batch_destroy(b);
                return fail_hip(e, "batch input allocation");
            }
        }
        std::vector<int32_t> idx(ng);
        std::vector<double> wx(p.nfreq), vc(nc);
        st = gpp_fill_seeded(&p, seeds[i], wx_center ? wx_center[i] : 4.0, dw ? dw[i] : 1.0, idx.data(), wx.data(),
                             vc.data(), nullptr, nullptr, nullptr, nullptr);
        hipError_t e = hipSuccess;
        if (!st) e = hipMemcpyAsync(o[0].ptr, idx.data(), bytes[0], hipMemcpyHostToDevice, s);
        if (!st && !e) e = hipMemcpyAsync(o[1].ptr, wx.data(), bytes[1], hipMemcpyHostToDevice, s);
        if (!st && !e) e = hipMemcpyAsync(o[2].ptr, vc.data(), bytes[2], hipMemcpyHostToDevice, s);
        if (!st && !e) e = gpp::launch_fill_seeded(static_cast<cplx *>(o[3].ptr), nb * nc, seeds[i], gpp::kArrAqsm, -1.0, 1.0, -1.0, 1.0, s);
        if (!st && !e) e = gpp::launch_fill_seeded(static_cast<cplx *>(o[4].ptr), nb * nc, seeds[i], gpp::kArrAqsn, -1.0, 1.0, -1.0, 1.0, s);
        if (!st && !e) e = gpp::launch_fill_seeded(static_cast<cplx *>(o[5].ptr), ng * nc, seeds[i], gpp::kArrIeps, -0.5, 0.5, -0.5, 0.5, s);
        if (!st && !e) e = gpp::launch_fill_seeded(static_cast<cplx *>(o[6].ptr), ng * nc, seeds[i], gpp::kArrWtilde, 0.5, 5.0, -0.05, 0.05, s);
        if (!st && !e) e = hipStreamSynchronize(s);  // the host vectors above go out of scope
        if (st || e) {
            gpp_batch_destroy(b);
            return st ? st : fail_hip(e, "batch input generation");
        }
        gpp::DevProblem &dp = b->dps[i];
        dp.wx_neg = gpp::wx_has_negative(wx.data(), p.nfreq);
        dp.idx = static_cast<const int32_t *>(o[0].ptr);
        dp.wx = static_cast<const double *>(o[1].ptr);
        dp.vcoul = static_cast<const double *>(o[2].ptr);
        dp.aqsm = static_cast<const cplx *>(o[3].ptr);
        dp.aqsn = static_cast<const cplx *>(o[4].ptr);
        dp.ieps = static_cast<const cplx *>(o[5].ptr);
        dp.wtilde = static_cast<const cplx *>(o[6].ptr);
    }
    *out = b;
    return GPP_OK;
}

gpp_status gpp_batch_sigma(gpp_batch *b, gpp_form form, const gpp_outputs *outs)
{
    if (!b || !outs) return fail(GPP_ERR_INVALID_ARGUMENT, "NULL argument");
    gpp::GramKind kind;
    gpp_status stk = batch_kind(b, form, &kind);
    if (stk) return stk;
    for (int i = 0; i < b->n; ++i) {
        if (!outs[i].achtemp || !outs[i].asxtemp)
            return fail(GPP_ERR_INVALID_ARGUMENT, "instance %d: achtemp and asxtemp are required", i);
        if (outs[i].acht_n1_loc)
            return fail(GPP_ERR_UNSUPPORTED, "instance %d: batches do not compute acht_n1_loc; pass NULL", i);
    }
    gpp_status st = set_device(b->ctx);
    if (st) return st;
    hipStream_t s = b->ctx->stream;
    if ((!b->planned || kind.ozaki_slices != b->kind.ozaki_slices || kind.int8 != b->kind.int8) &&
        (st = batch_prepare(b, kind)))
        return st;
    if (b->use_graph && !b->exec && (st = batch_capture(b))) return st;
    HIP_OR_FAIL(hipEventRecord(b->ev0, s), "hipEventRecord");
    if (b->use_graph) {
        HIP_OR_FAIL(hipGraphLaunch(b->exec, s), "hipGraphLaunch");
    } else {  // stage by stage, with events between the stages
        for (int st_i = 0; st_i < 4; ++st_i) {
            HIP_OR_FAIL(gpp::launch_batch(b->plan, b->work.ptr, batch_blocks(b), s, 1 << st_i), "batch launch");
            if (st_i < 3) HIP_OR_FAIL(hipEventRecord(b->evs[st_i], s), "hipEventRecord");
        }
    }
    HIP_OR_FAIL(hipEventRecord(b->ev1, s), "hipEventRecord");
    HIP_OR_FAIL(hipMemcpyAsync(b->host_out, static_cast<const char *>(b->work.ptr) + b->plan.tab_out,
                               b->plan.out_bytes, hipMemcpyDeviceToHost, s),
                "D2H batch results");
    HIP_OR_FAIL(hipStreamSynchronize(s), "batch synchronize");
    float ms = 0.f;
    HIP_OR_FAIL(hipEventElapsedTime(&ms, b->ev0, b->ev1), "hipEventElapsedTime");
    b->last_ms = ms;
    for (int q = 0; q < 4; ++q) b->stage_ms[q] = 0.0;
    if (!b->use_graph) {
        hipEvent_t marks[5] = {b->ev0, b->evs[0], b->evs[1], b->evs[2], b->ev1};
        for (int q = 0; q < 4; ++q) {
            HIP_OR_FAIL(hipEventElapsedTime(&ms, marks[q], marks[q + 1]), "hipEventElapsedTime");
            b->stage_ms[q] = ms;
        }
    }
    for (int i = 0; i < b->n; ++i) {
        const gpp::DevProblem &dp = b->dps[i];
        const cplx *r = reinterpret_cast<const cplx *>(reinterpret_cast<const char *>(b->host_out) +
                                                       (gpp::batch_results_offset(b->plan, i) - b->plan.tab_out));
        memcpy(outs[i].achtemp, r, sizeof(cplx) * dp.nfreq);
        memcpy(outs[i].asxtemp, r + dp.nfreq, sizeof(cplx) * dp.nfreq);
        if (outs[i].achstemp) memcpy(outs[i].achstemp, r + 2 * dp.nfreq + dp.nb, sizeof(cplx));
    }
    return GPP_OK;
}

gpp_status gpp_batch_autotune(gpp_batch *b, gpp_form form)
{
    if (!b) return fail(GPP_ERR_INVALID_ARGUMENT, "NULL batch");
    gpp::GramKind kind;
    gpp_status st = batch_kind(b, form, &kind);
    if (st) return st;
    st = set_device(b->ctx);
    if (st) return st;
    hipStream_t s = b->ctx->stream;
    // time the stages selected by `which` for kind k (best of two after a warm-up)
    auto timed = [&](const gpp::GramKind &k, int which, double *ms_out) -> gpp_status {
        gpp_status st2 = batch_prepare(b, k);
        if (st2) return st2;
        double best = 1e300;
        for (int rep = 0; rep < 3; ++rep) {
            HIP_OR_FAIL(gpp::launch_batch(b->plan, b->work.ptr, batch_blocks(b), s, 15 & ~which), "autotune launch");
            HIP_OR_FAIL(hipEventRecord(b->ev0, s), "hipEventRecord");
            HIP_OR_FAIL(gpp::launch_batch(b->plan, b->work.ptr, batch_blocks(b), s, which), "autotune launch");
            HIP_OR_FAIL(hipEventRecord(b->ev1, s), "hipEventRecord");
            HIP_OR_FAIL(hipStreamSynchronize(s), "autotune synchronize");
            float ms = 0.f;
            HIP_OR_FAIL(hipEventElapsedTime(&ms, b->ev0, b->ev1), "hipEventElapsedTime");
            if (rep) best = std::min(best, (double)ms);
        }
        *ms_out = best;
        return GPP_OK;
    };
    gpp::GramKind k = kind, best_k = kind;
    double best = 1e300, ms = 0;
    for (int tile = 0; tile < (kind.ozaki_slices ? 0 : gpp::kGramTiles); ++tile) {  // FP64 Gram tiles only
        k.tile = tile;
        if ((st = timed(k, 2, &ms))) return st;
        if (ms < best) best = ms, best_k.tile = tile;
    }
    k = best_k;
    best = 1e300;
    const int chunks[] = {2048, 4096, 8192, 16384};
    for (int chunk : chunks)
        for (int fc : {4, 8}) {
            k.red_chunk = chunk;
            k.red_fc = fc;
            if ((st = timed(k, 4, &ms))) return st;
            if (ms < best) best = ms, best_k.red_chunk = chunk, best_k.red_fc = fc;
        }
    return batch_prepare(b, best_k);
}

gpp_status gpp_batch_use_graph(gpp_batch *b, int use_graph)
{
    if (!b) return fail(GPP_ERR_INVALID_ARGUMENT, "NULL batch");
    b->use_graph = use_graph != 0;
    return GPP_OK;
}

gpp_status gpp_batch_last_timing(const gpp_batch *b, double *ms, int *gram_tile, int *reduce_chunk, int *reduce_freqs)
{
    if (!b) return fail(GPP_ERR_INVALID_ARGUMENT, "NULL batch");
    if (ms) *ms = b->last_ms;
    if (gram_tile) *gram_tile = b->kind.tile;
    if (reduce_chunk) *reduce_chunk = b->kind.red_chunk;
    if (reduce_freqs) *reduce_freqs = b->kind.red_fc;
    return GPP_OK;
}

gpp_status gpp_batch_last_stages(const gpp_batch *b, double *stage_ms)
{
    if (!b || !stage_ms) return fail(GPP_ERR_INVALID_ARGUMENT, "NULL argument");
    for (int q = 0; q < 4; ++q) stage_ms[q] = b->stage_ms[q];
    return GPP_OK;
}

gpp_status gpp_batch_destroy(gpp_batch *b)
{
    if (!b) return GPP_OK;
    if (b->ctx) {
        (void)hipSetDevice(b->ctx->device);
        (void)hipStreamSynchronize(b->ctx->stream);
    }
    batch_drop_graph(b);
    for (DevBuf &d : b->owned) d.release();
    b->work.release();
    if (b->host_out) (void)hipHostFree(b->host_out);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    for (hipEvent_t e : b->evs)
        if (e) (void)hipEventDestroy(e);
    delete b;
    return GPP_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Reference-compatible launchers.  They share one grow-only scratch area per
// device (partials, results, index flag), guarded by a mutex: calls from
// several host threads serialise on it.
namespace {

struct CompatScratch {
    gpp::Workspace work, results, flag;
};

std::mutex g_compat_mu;
std::vector<CompatScratch> g_compat;

gpp_status compat_run(int which, const gpp::DevProblem &dp, double *ach_re, double *ach_im, cplx *asx,
                      cplx *acht_n1, double *achs_re, double *achs_im)
{
    std::lock_guard<std::mutex> lock(g_compat_mu);
    int dev = 0;
    HIP_OR_FAIL(hipGetDevice(&dev), "hipGetDevice");
    if ((size_t)dev >= g_compat.size()) g_compat.resize(dev + 1);
    CompatScratch &cs = g_compat[dev];
    hipStream_t s = nullptr;  // null stream, as the reference launches
    HIP_OR_FAIL(cs.flag.reserve(sizeof(int)), "hipMalloc flag");
    HIP_OR_FAIL(hipMemsetAsync(cs.flag.ptr, 0, sizeof(int), s), "hipMemsetAsync");
    HIP_OR_FAIL(gpp::launch_check_index(dp.idx, dp.ngpown, static_cast<int *>(cs.flag.ptr), s), "check_index");
    int bad = 0;
    HIP_OR_FAIL(hipMemcpy(&bad, cs.flag.ptr, sizeof(int), hipMemcpyDeviceToHost), "D2H flag");
    if (bad) return fail(GPP_ERR_BAD_INDEX, "%d negative inv_igp_index entries", bad);
    HIP_OR_FAIL(cs.work.reserve(gpp::direct_workspace_bytes(dp)), "workspace allocation");
    gpp::DevResults res;
    res.nfreq = dp.nfreq;
    res.nb = dp.nb;
    HIP_OR_FAIL(cs.results.reserve(sizeof(cplx) * res.count()), "results allocation");
    res.base = static_cast<cplx *>(cs.results.ptr);
    HIP_OR_FAIL(gpp::run_direct_part(dp, which, cs.work.ptr, res, s), "direct launch");
    HIP_OR_FAIL(gpp::launch_accumulate_compat(res, which, dp.nv, ach_re, ach_im, asx, acht_n1, achs_re, achs_im, s),
                "accumulate launch");
    return GPP_OK;
}

gpp::DevProblem compat_problem(int nb, int nv, int ncouls, int ngpown, const int *idx, const double *wx,
                               const double *vcoul, const gpp_complex *aqsm, const gpp_complex *aqsn,
                               const gpp_complex *ieps, const gpp_complex *wtilde, double occ)
{
    gpp::DevProblem dp;
    dp.nb = nb; dp.nv = nv; dp.ncouls = ncouls; dp.ngpown = ngpown; dp.nfreq = 3; dp.occ = occ;
    dp.idx = idx; dp.wx = wx; dp.vcoul = vcoul;
    dp.aqsm = reinterpret_cast<const cplx *>(aqsm);
    dp.aqsn = reinterpret_cast<const cplx *>(aqsn);
    dp.ieps = reinterpret_cast<const cplx *>(ieps);
    dp.wtilde = reinterpret_cast<const cplx *>(wtilde);
    return dp;
}

}  // namespace

extern "C" {

gpp_status d_till_nvband(int nvband, int ngpown, int ncouls, int *inv_igp_index, double *wx_array,
                         gpp_complex *wtilde_array, gpp_complex *aqsmtemp, gpp_complex *aqsntemp,
                         gpp_complex *I_eps_array, gpp_complex *asxtemp, double *vcoul,
                         gpp_complex *acht_n1_loc, double *achtemp_re, double *achtemp_im, const double occ)
{
    if (nvband < 0 || ngpown < 1 || ncouls < 1)
        return fail(GPP_ERR_INVALID_ARGUMENT, "invalid sizes nvband=%d ngpown=%d ncouls=%d", nvband, ngpown, ncouls);
    if (nvband == 0) return GPP_OK;
    if (!inv_igp_index || !wx_array || !wtilde_array || !aqsmtemp || !aqsntemp || !I_eps_array || !vcoul ||
        !achtemp_re || !achtemp_im)
        return fail(GPP_ERR_INVALID_ARGUMENT, "NULL device array");
    gpp::DevProblem dp = compat_problem(nvband, nvband, ncouls, ngpown, inv_igp_index, wx_array, vcoul, aqsmtemp,
                                        aqsntemp, I_eps_array, wtilde_array, occ);
    return compat_run(0, dp, achtemp_re, achtemp_im, reinterpret_cast<cplx *>(asxtemp),
                      reinterpret_cast<cplx *>(acht_n1_loc), nullptr, nullptr);
}

gpp_status d_reduce_achstemp(int number_bands, int *inv_igp_index, int ncouls, gpp_complex *aqsmtemp,
                             gpp_complex *aqsntemp, gpp_complex *I_eps_array, double *achstemp_re,
                             double *achstemp_im, int ngpown, double *vcoul)
{
    if (number_bands < 1 || ngpown < 1 || ncouls < 1)
        return fail(GPP_ERR_INVALID_ARGUMENT, "invalid sizes number_bands=%d ngpown=%d ncouls=%d", number_bands,
                    ngpown, ncouls);
    if (!inv_igp_index || !aqsmtemp || !aqsntemp || !I_eps_array || !vcoul || !achstemp_re || !achstemp_im)
        return fail(GPP_ERR_INVALID_ARGUMENT, "NULL device array");
    gpp::DevProblem dp = compat_problem(number_bands, 0, ncouls, ngpown, inv_igp_index, nullptr, vcoul, aqsmtemp,
                                        aqsntemp, I_eps_array, nullptr, 1.0);
    return compat_run(2, dp, nullptr, nullptr, nullptr, nullptr, achstemp_re, achstemp_im);
}

gpp_status d_noFlagOCCSolver(int number_bands, int nvband, int ngpown, int ncouls, int *inv_igp_index,
                             double *wx_array, gpp_complex *wtilde_array, gpp_complex *aqsmtemp,
                             gpp_complex *aqsntemp, gpp_complex *I_eps_array, double *vcoul, double *achtemp_re,
                             double *achtemp_im, gpp_complex *acht_n1_loc)
{
    if (number_bands < 1 || nvband < 0 || nvband > number_bands || ngpown < 1 || ncouls < 1)
        return fail(GPP_ERR_INVALID_ARGUMENT, "invalid sizes number_bands=%d nvband=%d ngpown=%d ncouls=%d",
                    number_bands, nvband, ngpown, ncouls);
    if (nvband == number_bands) return GPP_OK;
    if (!inv_igp_index || !wx_array || !wtilde_array || !aqsmtemp || !aqsntemp || !I_eps_array || !vcoul ||
        !achtemp_re || !achtemp_im)
        return fail(GPP_ERR_INVALID_ARGUMENT, "NULL device array");
    gpp::DevProblem dp = compat_problem(number_bands, nvband, ncouls, ngpown, inv_igp_index, wx_array, vcoul,
                                        aqsmtemp, aqsntemp, I_eps_array, wtilde_array, 1.0);
    return compat_run(1, dp, achtemp_re, achtemp_im, nullptr, reinterpret_cast<cplx *>(acht_n1_loc), nullptr,
                      nullptr);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// The launcher that allocates nothing: d_noFlagOCCSolver with its scratch --
// the index flag, the results and the direct form's partial sums -- in one
// file-scope device array instead of the grow-only per-device workspace.
namespace {

constexpr size_t kNoAllocScratchBytes = (size_t)16 << 20;  // 16 MiB per device
constexpr int kNoAllocMaxDevices = 64;

}  // namespace

// The scratch array.  In host code this name is only a stand-in for the
// symbol: its device address (one per device) comes from hipGetSymbolAddress
// and is the only handle host code passes on.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

void *g_noalloc_ptr[kNoAllocMaxDevices] = {};  // device address of scratch on each device (guarded by g_compat_mu)

// Device address of the scratch array on the current device `dev`, looked up
// on the first call there.
gpp_status noalloc_scratch(int dev, char **base)
{
    if (dev < 0 || dev >= kNoAllocMaxDevices)
        return fail(GPP_ERR_NO_DEVICE, "device %d beyond the %d the no-alloc launcher serves", dev, kNoAllocMaxDevices);
    if (!g_noalloc_ptr[dev]) {
        void *ptr = nullptr;
        HIP_OR_FAIL(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!ptr) return fail(GPP_ERR_LAUNCH, "hipGetSymbolAddress(scratch) returned NULL");
        g_noalloc_ptr[dev] = ptr;
    }
    *base = static_cast<char *>(g_noalloc_ptr[dev]);
    return GPP_OK;
}

}  // namespace

extern "C" {

gpp_status d_noFlagOCCSolver_noalloc(int number_bands, int nvband, int ngpown, int ncouls, int *inv_igp_index,
                                     double *wx_array, gpp_complex *wtilde_array, gpp_complex *aqsmtemp,
                                     gpp_complex *aqsntemp, gpp_complex *I_eps_array, double *vcoul,
                                     double *achtemp_re, double *achtemp_im, gpp_complex *acht_n1_loc)
{
    if (number_bands < 1 || nvband < 0 || nvband > number_bands || ngpown < 1 || ncouls < 1)
        return fail(GPP_ERR_INVALID_ARGUMENT, "invalid sizes number_bands=%d nvband=%d ngpown=%d ncouls=%d",
                    number_bands, nvband, ngpown, ncouls);
    if (nvband == number_bands) return GPP_OK;
    if (!inv_igp_index || !wx_array || !wtilde_array || !aqsmtemp || !aqsntemp || !I_eps_array || !vcoul ||
        !achtemp_re || !achtemp_im)
        return fail(GPP_ERR_INVALID_ARGUMENT, "NULL device array");
    const gpp::DevProblem dp = compat_problem(number_bands, nvband, ncouls, ngpown, inv_igp_index, wx_array, vcoul,
                                              aqsmtemp, aqsntemp, I_eps_array, wtilde_array, 1.0);
    // scratch: [index flag: 256 B][results][direct-form partial sums]
    gpp::DevResults res;
    res.nfreq = dp.nfreq;
    res.nb = dp.nb;
    const size_t res_bytes = (sizeof(cplx) * res.count() + 255) & ~(size_t)255;
    const size_t need = 256 + res_bytes + gpp::direct_workspace_bytes(dp);
    if (need > kNoAllocScratchBytes)
        return fail(GPP_ERR_INVALID_ARGUMENT,
                    "number_bands=%d ngpown=%d needs %zu bytes of scratch; d_noFlagOCCSolver_noalloc holds %zu "
                    "(use d_noFlagOCCSolver)",
                    number_bands, ngpown, need, kNoAllocScratchBytes);
    std::lock_guard<std::mutex> lock(g_compat_mu);  // calls share the device's scratch: one at a time
    int dev = 0;
    HIP_OR_FAIL(hipGetDevice(&dev), "hipGetDevice");
    char *base = nullptr;
    gpp_status st = noalloc_scratch(dev, &base);
    if (st) return st;
    hipStream_t s = nullptr;  // null stream, as the reference launches
    int *flag = reinterpret_cast<int *>(base);
    HIP_OR_FAIL(gpp::launch_zero_ints(flag, 1, s), "zero flag");  // the last call's flag stays in scratch
    HIP_OR_FAIL(gpp::launch_check_index(dp.idx, dp.ngpown, flag, s), "check_index");
    int bad = 0;
    HIP_OR_FAIL(hipMemcpy(&bad, flag, sizeof(int), hipMemcpyDeviceToHost), "D2H flag");
    if (bad) return fail(GPP_ERR_BAD_INDEX, "%d negative inv_igp_index entries", bad);
    res.base = reinterpret_cast<cplx *>(base + 256);
    HIP_OR_FAIL(gpp::run_direct_part(dp, 1, base + 256 + res_bytes, res, s), "direct launch");
    HIP_OR_FAIL(gpp::launch_accumulate_compat(res, 1, dp.nv, achtemp_re, achtemp_im, nullptr,
                                              reinterpret_cast<cplx *>(acht_n1_loc), nullptr, nullptr, s),
                "accumulate launch");
    return GPP_OK;
}

}  // extern "C"
