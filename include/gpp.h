/*
 * gpp.h -- C ABI of the MI355X (gfx950) GPP sigma library, libgpp_hip.so.
 *
 * The library evaluates the band / G-vector loop of the BerkeleyGW GPP
 * (General Plasmon Pole) sigma mini-app on one GPU with hand-written HIP
 * kernels.  It replaces two seams of the reference mini-app:
 *
 *   1. the timed region of gppComplex.cpp:332-387 (main), which a host
 *      driver replaces by gpp_upload() + gpp_sigma();
 *   2. the device launchers declared at GPPComplex.h:306-310
 *      (d_till_nvband, d_reduce_achstemp, d_noFlagOCCSolver), which take
 *      caller-owned device pointers and accumulate into caller-zeroed device
 *      outputs, as gppCudaVer.cpp:389-392 calls them.
 *
 * Conventions
 *   - Complex values are gpp_complex {double re, im}: 16 bytes, interleaved,
 *     layout-compatible with the reference's GPPComplex (Complex.h:15-19),
 *     std::complex<double> and hipDoubleComplex.
 *   - Arrays are row-major: aqsmtemp/aqsntemp [number_bands][ncouls],
 *     I_eps_array/wtilde_array [ngpown][ncouls], vcoul [ncouls],
 *     inv_igp_index [ngpown] (int32), wx_array [nfreq].
 *   - igp = min(inv_igp_index[j], ncouls-1) as gppComplex.cpp:342-344; a
 *     negative entry is rejected with GPP_ERR_BAD_INDEX (the reference reads
 *     out of bounds there).
 *   - Every entry point returns a gpp_status; nothing prints or exits.
 *     gpp_last_error() gives a message for the calling thread's last error.
 *   - Results are deterministic: the same inputs give bitwise-identical
 *     outputs on every call (fixed-order reductions, no floating atomics).
 *   - One context per host thread; a context is not re-entrant.
 */
#ifndef GPP_H
#define GPP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GPP_ABI_VERSION 1

typedef enum {
    GPP_OK = 0,
    GPP_ERR_INVALID_ARGUMENT = 1, /* bad size, null pointer, nvband > number_bands ... */
    GPP_ERR_BAD_INDEX = 2,        /* negative inv_igp_index entry */
    GPP_ERR_NO_DEVICE = 3,        /* no usable gfx950 device, or HIP runtime error at init */
    GPP_ERR_OUT_OF_MEMORY = 4,    /* device allocation failed */
    GPP_ERR_LAUNCH = 5,           /* kernel launch or execution error */
    GPP_ERR_UNSUPPORTED = 6,      /* form not available for this problem */
    GPP_ERR_NO_INPUTS = 7,        /* gpp_sigma before gpp_upload / gpp_bind_device */
    GPP_ERR_NONFINITE = 8,        /* reserved: non-finite output reported by a check */
    GPP_ERR_IO = 9                /* problem file: open/read/write failure, bad header, truncation, checksum */
} gpp_status;

/* One complex number, {re, im}; Complex.h:15-19. */
typedef struct {
    double re;
    double im;
} gpp_complex;

/*
 * Problem sizes.  nfreq frequencies wx_array[0..nfreq) (the reference fixes
 * nfreq = nend - nstart = 3, gppComplex.cpp:12-13).  occ multiplies asxtemp
 * (gppComplex.cpp:247, :363).
 */
typedef struct {
    int32_t number_bands;
    int32_t nvband;
    int32_t ncouls;
    int32_t ngpown;
    int32_t nfreq;
    double occ;
} gpp_problem;

/* Evaluation forms of the band loop. */
typedef enum {
    /* Direct FP64 kernel over (n1, igp, ig, iw): each update evaluates its own
     * pole factor, complex division included. */
    GPP_FORM_DIRECT = 0,
    /* Direct FP64 kernel evaluating each term in the reference's exact
     * operation order (IEEE division, no FMA contraction). */
    GPP_FORM_DIRECT_STRICT = 1,
    /* Band sum factored into the complex Gram product
     * M[ig,j] = sum_n1 aqsn[n1,ig] conj(aqsm[n1,igp(j)]) on FP64 MFMA,
     * followed by a fused FP64 pole-factor reduce. */
    GPP_FORM_FACTORED_F64 = 2,
    /* As GPP_FORM_FACTORED_F64 with the Gram product split into FP16 slices
     * (Ozaki scheme: digits in base 254 -- base 128 at 6 slices -- under a
     * power-of-two scale per column), multiplied on FP16 MFMA with FP32
     * accumulation (exact for the leading digit levels; the rest round below
     * the split's truncation) and recombined in FP64; slices per operand from
     * GPP_OPT_OZAKI_SLICES. */
    GPP_FORM_FACTORED_OZAKI = 3,
    /* As GPP_FORM_FACTORED_OZAKI with the digits (base 254 at every slice
     * count) held in INT8 and multiplied on INT8 MFMA with exact INT32
     * accumulation (flushed to FP64 every 8192 bands). */
    GPP_FORM_FACTORED_OZAKI_I8 = 4
} gpp_form;

/* Tunable options of a context (gpp_ctx_set_option / gpp_ctx_get_option). */
typedef enum {
    /* Slices per operand of the Ozaki forms, 3..6 (default 6).  The first
     * slice carries 7 bits, each later one about 8 (base 254): the split
     * truncates at about 2^-39 of the column scale at 5 slices and 2^-47 at
     * 6 (FP16 at 6 slices: 7-bit digits, 2^-42).  5 slices run 15 digit
     * products instead of 21 (about 25 % faster) for problems whose sums do
     * not cancel much (2.4e-11 at 512/2/32768/1638; 3.1e-10 at
     * 1536/8/65536/3276 with 16 frequencies, where 6 are needed). */
    GPP_OPT_OZAKI_SLICES = 1,
    /* Block layout of the Ozaki Gram kernel (all give bit-identical results):
     * 0 = 64x64 tiles, 2x2 waves of 32x32 (one wave per SIMD); 1 = 64x64
     * tiles, 2x4 waves of 32x16 (two per SIMD), both on 16x16 MFMAs staging
     * each band step through registers; 2 = 64x32 tiles, 2x2 waves of 32x16
     * on 16x16 MFMAs, band steps loaded straight into LDS, two blocks per
     * compute unit; 3 = 64x64 tiles, 2x2 waves of 32x32 on 32x32 MFMAs, band
     * steps loaded straight into two to four LDS buffers ahead of use; 4 =
     * layout 3 in persistent blocks, one per compute unit, whose tiles form
     * one stream of band steps (default). */
    GPP_OPT_OZAKI_LAYOUT = 2,
    /* FP64-MFMA Gram tile (G' columns x G-vectors): 0 = 128x128 (default),
     * 1 = 64x128, 2 = 128x64, 3 = 64x64. */
    GPP_OPT_GRAM_TILE = 3,
    /* G-vectors per pole-reduce block: a power of two in [256, 2^20]
     * (default 8192). */
    GPP_OPT_REDUCE_CHUNK = 4,
    /* Frequencies per pole-reduce pass when nfreq > 4: 4, 8 or 16; 0 = 16
     * from 16 frequencies up when none is negative, else 8 (default).  The
     * last passes are narrower: 8 then 4 wide with 16 (24 frequencies run
     * 16 + 8), one of 8, 6, 4 or 2 with 8 (13 run 8 + 6).  Negative
     * frequencies run at most 8.  The
     * batched form runs 4 for 4, 16 for 16 and 8 otherwise (16-frequency
     * passes timed slower on its uneven work lists, 11.9 against 11.5 ms). */
    GPP_OPT_REDUCE_FREQS = 5,
    /* With band-dependent frequencies (gpp_set_band_frequencies): the most
     * groups of bands sharing one frequency row that the factored forms
     * factor group by group; beyond it they run the direct form (default
     * 32). */
    GPP_OPT_MAX_BUCKETS = 6,
    /* Frequencies one pole-reduce block evaluates, in passes of
     * GPP_OPT_REDUCE_FREQS, re-reading its G-vector range from the caches
     * between passes: 0 = all nfreq (default); k > 0 splits the frequencies
     * over ceil(nfreq / k) blocks per range. */
    GPP_OPT_REDUCE_FREQS_PER_BLOCK = 7,
    /* G' column tiles per tile group of the Ozaki Gram kernel: the blocks of
     * a group share those tiles' digit planes in L2 and stream the G-vector
     * digit planes once per group.  0 = sized for the L2 (default). */
    GPP_OPT_OZAKI_GROUP = 8,
    /* GPP_FORM_FACTORED_F64 with at most 3 frequencies, none negative, at
     * most 4 occupied bands, the 128x128 Gram tile and acht_n1_loc not
     * requested: 1 = evaluate the pole factors in the epilogue of the Gram
     * product, so the G'xG product Mu (ngpown * ncouls * 16 bytes of
     * workspace) is never stored -- for problems whose Mu does not fit in
     * device memory; slower where it fits (512/2/32768/1638: 3.58 against
     * 3.19 ms); 0 = store Mu and run the pole reduce as its own stage
     * (default). */
    GPP_OPT_FUSED_REDUCE = 9
} gpp_option;

/* Outputs of one sigma evaluation (caller-owned host memory, overwritten).
 * achtemp and asxtemp have nfreq entries; acht_n1_loc has number_bands
 * entries (frequency 2 of each band, as gppComplex.cpp:381; frequency nfreq-1
 * when nfreq < 3) and may be NULL; achstemp (reduce_achstemp,
 * gppComplex.cpp:66-101) may be NULL. */
typedef struct {
    gpp_complex *achtemp;
    gpp_complex *asxtemp;
    gpp_complex *acht_n1_loc;
    gpp_complex *achstemp;
} gpp_outputs;

typedef struct gpp_ctx gpp_ctx;

/* ---- library / context ---------------------------------------------------- */

/* Message for the calling thread's most recent failed call ("" if none). */
const char *gpp_last_error(void);

/* ABI version compiled into the library (GPP_ABI_VERSION). */
int gpp_abi_version(void);

/* Creates a context on HIP device `device`: a default-priority stream and
 * timing events.  No input buffers are allocated until gpp_upload. */
gpp_status gpp_ctx_create(gpp_ctx **out, int device);

/* Releases the context and every device buffer it owns. Null is a no-op. */
gpp_status gpp_ctx_destroy(gpp_ctx *ctx);

/* ---- inputs ----------------------------------------------------------------- */

/*
 * Copies one problem's inputs from host memory into context-owned device
 * buffers (synchronous; untimed, as gppCudaVer.cpp:368-377).  Buffers are
 * reused across uploads of equal or smaller size.  Validates sizes and
 * inv_igp_index.
 */
gpp_status gpp_upload(gpp_ctx *ctx, const gpp_problem *prob, const int32_t *inv_igp_index,
                      const double *wx_array, const double *vcoul, const gpp_complex *aqsmtemp,
                      const gpp_complex *aqsntemp, const gpp_complex *I_eps_array,
                      const gpp_complex *wtilde_array);

/*
 * Points the context at caller-owned DEVICE buffers with the same layout
 * (no copy).  The caller keeps them alive and unchanged while the context
 * uses them.  inv_igp_index is validated on the device.
 */
gpp_status gpp_bind_device(gpp_ctx *ctx, const gpp_problem *prob, const int32_t *d_inv_igp_index,
                           const double *d_wx_array, const double *d_vcoul,
                           const gpp_complex *d_aqsmtemp, const gpp_complex *d_aqsntemp,
                           const gpp_complex *d_I_eps_array, const gpp_complex *d_wtilde_array);

/*
 * Fills host arrays with the reference's deterministic inputs,
 * gppComplex.cpp:296-329: aqsmtemp = aqsntemp = (0.5+0.5i)(n1+ig),
 * I_eps = wtilde = (0.5+0.5i)(j+ig), inv_igp_index[j] = (j+1)*ncouls/ngpown
 * (igp_zero_based != 0 gives j*ncouls/ngpown), vcoul[ig] = ig,
 * wx[iw] = e_lk - e_n1kq + dw*(iw-1) clamped to >= 1e-6 with e_lk=10,
 * e_n1kq=6, dw=1.  Any output pointer may be NULL to skip that array.
 */
gpp_status gpp_fill_deterministic(const gpp_problem *prob, int igp_zero_based,
                                  int32_t *inv_igp_index, double *wx_array, double *vcoul,
                                  gpp_complex *aqsmtemp, gpp_complex *aqsntemp,
                                  gpp_complex *I_eps_array, gpp_complex *wtilde_array);

/*
 * Fills host arrays with seeded pseudo-random inputs: aqsm/aqsn re,im ~
 * U(-1,1); I_eps re,im ~ U(-0.5,0.5); wtilde re ~ U(0.5,5), im ~
 * U(-0.05,0.05); vcoul ~ U(0.1,10); inv_igp_index = sorted distinct random
 * G-vector indices.  wx_array is filled as wx_center + dw*(iw - (nfreq-1)/2)
 * (wx_center=4, dw=1 gives the reference's 3,4,5).  Element values are a
 * counter-based hash of (seed, array, index), so any sub-range can be
 * regenerated independently (gpp_fill_seeded_device produces the same bits).
 */
gpp_status gpp_fill_seeded(const gpp_problem *prob, uint64_t seed, double wx_center, double dw,
                           int32_t *inv_igp_index, double *wx_array, double *vcoul,
                           gpp_complex *aqsmtemp, gpp_complex *aqsntemp,
                           gpp_complex *I_eps_array, gpp_complex *wtilde_array);

/*
 * Generates the seeded inputs of gpp_fill_seeded directly into the
 * context's own device buffers (no host copy of the large arrays; the small
 * arrays are generated on the host and copied).  Leaves the context ready
 * for gpp_sigma.
 */
gpp_status gpp_generate_seeded(gpp_ctx *ctx, const gpp_problem *prob, uint64_t seed,
                               double wx_center, double dw);

/* ---- evaluation ------------------------------------------------------------- */

/*
 * Evaluates achtemp/asxtemp (and the optional outputs) for the inputs last
 * uploaded or bound, with the given form.  Synchronous.  The GPU time of the
 * call is available from gpp_ctx_last_timing.  With the factored forms,
 * acht_n1_loc costs a second product of the size of the Gram product (FP64
 * MFMA whatever the form) and more workspace (16 + 1 bytes per
 * ngpown x ncouls element); pass NULL when it is not needed.
 */
gpp_status gpp_sigma(gpp_ctx *ctx, gpp_form form, const gpp_outputs *out);

/*
 * Band-dependent frequencies: band n1 uses wx_bands[n1 * nfreq + iw]
 * (host memory, [number_bands][nfreq], copied) instead of the shared
 * wx_array, as BerkeleyGW's wx = e_lk - e_n1kq(n1) + dw*(iw-1); the mini-app
 * fixes one e_n1kq for all bands (gppComplex.cpp:326), so its wx_array is
 * band-independent.  Applies to the inputs last uploaded or bound, until the
 * next gpp_upload / gpp_bind_device / gpp_generate_seeded; NULL restores the
 * shared wx_array.  The direct forms read the band's row.  The band sum
 * factors only over bands sharing a row, so the factored forms group the
 * bands by identical rows and factor each group (one group: exactly the
 * plain factored form on that row); with more than GPP_OPT_MAX_BUCKETS
 * groups they run GPP_FORM_DIRECT instead (gpp_ctx_last_form tells which).
 * The batched entry points take one wx_array per instance.
 */
gpp_status gpp_set_band_frequencies(gpp_ctx *ctx, const double *wx_bands);

/* Number of distinct frequency rows of the current inputs (1 when shared). */
gpp_status gpp_ctx_band_buckets(const gpp_ctx *ctx, int *nbuckets);

/* The form the last gpp_sigma on this context actually ran. */
gpp_status gpp_ctx_last_form(const gpp_ctx *ctx, gpp_form *form);

/* ---- problem files (.gpp) ----------------------------------------------------
 *
 * A problem file holds one problem's inputs and, optionally, its expected
 * outputs, so that inputs produced elsewhere (BerkeleyGW's eps / wtilde /
 * vcoul, or golden vectors) go through the same ABI.  Layout, little-endian:
 * a 128-byte header (magic "GPPSIG01", version 1, flags, the gpp_problem
 * sizes, occ, payload size, payload checksum), then the sections
 * inv_igp_index int32[ngpown], wx double[nfreq] (flag GPP_FILE_WX_PER_BAND:
 * [number_bands][nfreq]), vcoul double[ncouls], aqsmtemp, aqsntemp
 * complex[number_bands][ncouls], I_eps_array, wtilde_array
 * complex[ngpown][ncouls] and (flag GPP_FILE_HAS_EXPECTED) achtemp[nfreq],
 * asxtemp[nfreq], acht_n1_loc[number_bands], achstemp[1], each section
 * starting at a multiple of 64 bytes of the payload with zero padding
 * between.  Checksum: FNV-1a over the payload's 64-bit little-endian words
 * in four interleaved lanes: lane k (k = 0..3) starts at 0xcbf29ce484222325
 * and takes the words w_i with i % 4 == k, h_k = (h_k ^ w_i) * 0x100000001b3
 * mod 2^64; then h = 0xcbf29ce484222325 and h = (h ^ h_k) * 0x100000001b3 for
 * k = 0..3.
 */
#define GPP_FILE_WX_PER_BAND 1u
#define GPP_FILE_HAS_EXPECTED 2u

typedef struct {
    gpp_problem problem;
    uint32_t flags;          /* GPP_FILE_* */
    uint64_t payload_bytes;  /* bytes after the 128-byte header */
    uint64_t checksum;
} gpp_file_header;

/* Writes a problem file.  wx holds nfreq values, or number_bands x nfreq
 * when wx_per_band is non-zero.  expected may be NULL; if given, achtemp and
 * asxtemp are required and a NULL acht_n1_loc / achstemp is stored as NaN. */
gpp_status gpp_file_write(const char *path, const gpp_problem *prob, const int32_t *inv_igp_index,
                          const double *wx, int wx_per_band, const double *vcoul,
                          const gpp_complex *aqsmtemp, const gpp_complex *aqsntemp,
                          const gpp_complex *I_eps_array, const gpp_complex *wtilde_array,
                          const gpp_outputs *expected);

/* Reads and validates a file's header (the payload is not read). */
gpp_status gpp_file_info(const char *path, gpp_file_header *info);

/* Reads a whole file into caller-allocated host arrays sized from
 * gpp_file_info (wx: number_bands x nfreq with GPP_FILE_WX_PER_BAND) and
 * verifies the checksum.  expected may be NULL; it is filled only when the
 * file holds expected outputs (its NULL members are skipped). */
gpp_status gpp_file_read(const char *path, int32_t *inv_igp_index, double *wx, double *vcoul,
                         gpp_complex *aqsmtemp, gpp_complex *aqsntemp, gpp_complex *I_eps_array,
                         gpp_complex *wtilde_array, gpp_outputs *expected);

/* Uploads a problem file into the context as gpp_upload does, streaming the
 * large sections from disk to the device through two pinned 32 MB staging
 * buffers (host memory use does not grow with the problem), and sets the
 * band frequencies when the file holds them.  The checksum is verified
 * before the context accepts the inputs.  expected as for gpp_file_read. */
gpp_status gpp_upload_file(gpp_ctx *ctx, const char *path, gpp_outputs *expected);

/*
 * Enqueues one evaluation on the context's stream and returns without
 * waiting: evaluations enqueued back to back run without host round trips
 * (as the reference's CUDA driver launches and synchronises once,
 * gppCudaVer.cpp:381-395).  The results of the last one stay on the device
 * until gpp_ctx_wait; with_acht_n1_loc asks the factored forms for
 * acht_n1_loc (the direct forms always compute it).
 */
gpp_status gpp_sigma_async(gpp_ctx *ctx, gpp_form form, int with_acht_n1_loc);

/* Waits for the evaluations enqueued on the context and copies the last one's
 * outputs into out (may be NULL: only wait; NULL members are skipped). */
gpp_status gpp_ctx_wait(gpp_ctx *ctx, const gpp_outputs *out);

/*
 * Enqueues the context's work on the caller's HIP stream (hipStream_t of the
 * context's device) instead of its own: uploads, evaluations, timing events,
 * result copies and batch replays.  Work the caller enqueued on that stream
 * before gpp_sigma_async runs before the evaluation, and the caller's later
 * work runs after it, so a host application that owns its streams orders the
 * library's work with its own without synchronising (the reference's
 * launchers run on the default stream and leave synchronisation to the
 * caller, gppCudaVer.cpp:381-395).  The context first waits for the work
 * already enqueued where it ran so far.  NULL returns to the context's own
 * stream.  The caller keeps the stream alive until then or until
 * gpp_ctx_destroy.
 */
struct ihipStream_t;
gpp_status gpp_ctx_set_stream(gpp_ctx *ctx, struct ihipStream_t *stream);

/*
 * Restricts the following evaluations to the G' columns [j0, j1) of the
 * current inputs (rows j0..j1-1 of I_eps_array / wtilde_array and entries of
 * inv_igp_index): the slab one rank owns in the reference's MPI build
 * (gppMPIOpenMP3.cpp:172, 255-271).  The outputs are that slab's partial
 * sums (achtemp, asxtemp, acht_n1_loc and achstemp are sums over columns);
 * adding the slabs' outputs in rank order gives the whole problem's, as the
 * reference's MPI_Reduce does (gppMPIOpenMP3.cpp:361-370), up to the order
 * of the additions.  j0 = j1 = 0 restores every column; new inputs reset it.
 */
gpp_status gpp_ctx_set_columns(gpp_ctx *ctx, int j0, int j1);

/* Sets / reads a context option (gpp_option). */
gpp_status gpp_ctx_set_option(gpp_ctx *ctx, gpp_option option, int value);
gpp_status gpp_ctx_get_option(const gpp_ctx *ctx, gpp_option option, int *value);

/*
 * GPU time of the last gpp_sigma in milliseconds, measured with HIP events
 * on the context's stream: total, and per stage (stage_ms[i] for i <
 * *nstages, at most max_stages written; stage names via
 * gpp_ctx_stage_name).  Any pointer may be NULL.
 */
gpp_status gpp_ctx_last_timing(const gpp_ctx *ctx, double *total_ms, double *stage_ms,
                               int max_stages, int *nstages);

/*
 * Average GPU time, in milliseconds, of up to the last last_n evaluations
 * (at most 32 are kept) that ran the same stages as the last one: *nevals
 * evaluations averaged, total and per stage as gpp_ctx_last_timing.  Any
 * output pointer may be NULL.
 */
gpp_status gpp_ctx_timing_summary(const gpp_ctx *ctx, int last_n, int *nevals, double *avg_total_ms,
                                  double *avg_stage_ms, int max_stages, int *nstages);

/* Name of stage i of the last gpp_sigma ("" when out of range). */
const char *gpp_ctx_stage_name(const gpp_ctx *ctx, int i);

/*
 * Times the candidate Gram tiles and pole-reduce schedules of a factored
 * form on the inputs last uploaded or bound, and sets the fastest as the
 * context's GPP_OPT_GRAM_TILE / GPP_OPT_REDUCE_CHUNK / GPP_OPT_REDUCE_FREQS.
 * Results are cached per problem shape in the context.  (GPP_FORM_DIRECT
 * forms have nothing to tune and return GPP_OK.)
 */
gpp_status gpp_ctx_autotune(gpp_ctx *ctx, gpp_form form);

/* ---- batches: many independent problems in one pass ------------------------- */
/*
 * A batch evaluates n independent problems of different sizes with the
 * factored FP64 form: every stage is one persistent launch over a work list
 * spanning all instances, and the launches are captured once into a hipGraph
 * that each gpp_batch_sigma replays.  The batch uses its context's device
 * and stream; destroy it before the context.
 */
typedef struct gpp_batch gpp_batch;

/* Device arrays of one problem (same layout as gpp_bind_device). */
typedef struct {
    const int32_t *inv_igp_index;
    const double *wx_array;
    const double *vcoul;
    const gpp_complex *aqsmtemp;
    const gpp_complex *aqsntemp;
    const gpp_complex *I_eps_array;
    const gpp_complex *wtilde_array;
} gpp_device_inputs;

/* Batch over caller-owned device inputs (kept alive and unchanged). */
gpp_status gpp_batch_create(gpp_ctx *ctx, int n, const gpp_problem *probs, const gpp_device_inputs *inputs,
                            gpp_batch **out);

/* Batch over library-owned inputs generated on the device like
 * gpp_generate_seeded, instance i with seeds[i] and the frequency grid
 * wx_center[i] + dw[i]*(iw - (nfreq-1)/2) (NULL arrays: 4.0 and 1.0). */
gpp_status gpp_batch_create_seeded(gpp_ctx *ctx, int n, const gpp_problem *probs, const uint64_t *seeds,
                                   const double *wx_center, const double *dw, gpp_batch **out);

/* Evaluates every instance; outs[i] receives instance i (acht_n1_loc must be
 * NULL).  The factored forms are batched: GPP_FORM_FACTORED_F64,
 * GPP_FORM_FACTORED_OZAKI and GPP_FORM_FACTORED_OZAKI_I8 (with the context's
 * GPP_OPT_OZAKI_SLICES).  A form other than the last one re-plans the workspace and
 * re-captures the graph; results are bitwise equal to the single-problem
 * path of the same form. */
gpp_status gpp_batch_sigma(gpp_batch *b, gpp_form form, const gpp_outputs *outs);

/* Times the Gram tiles (FP64 form) and reduce schedules of `form` on this
 * batch and keeps the fastest (re-planning and re-capturing the graph). */
gpp_status gpp_batch_autotune(gpp_batch *b, gpp_form form);

/* 1 (default): replay the captured hipGraph; 0: launch the same kernels eagerly. */
gpp_status gpp_batch_use_graph(gpp_batch *b, int use_graph);

/* GPU milliseconds of the last gpp_batch_sigma (HIP events around the
 * launches) and the configuration in use (any pointer may be NULL). */
gpp_status gpp_batch_last_timing(const gpp_batch *b, double *ms, int *gram_tile, int *reduce_chunk,
                                 int *reduce_freqs);

/* Per-stage GPU milliseconds of the last gpp_batch_sigma run without the
 * graph (gpp_batch_use_graph(b, 0)): [gather + achstemp, Gram product,
 * pole reduce, final]; zeros after a graph replay. */
gpp_status gpp_batch_last_stages(const gpp_batch *b, double *stage_ms);

gpp_status gpp_batch_destroy(gpp_batch *b);

/* ---- reference-compatible device launchers (GPPComplex.h:306-310) ---------- */
/*
 * Same parameter lists and meaning as the reference's CUDA launchers, with
 * GPPComplex* replaced by the layout-identical gpp_complex*.  All pointers
 * are DEVICE pointers owned by the caller; nfreq is fixed at 3 (nstart=0,
 * nend=3).  Results ACCUMULATE (+=) into the caller-zeroed device outputs,
 * as gppCudaVer.cpp:296-298,376-377 expects.  Work is enqueued on the null
 * stream; the caller synchronises (gppCudaVer.cpp:395).  Unlike the
 * reference these return a status, validate inv_igp_index, and are
 * race-free and deterministic; asxtemp is accumulated (the reference never
 * zeroes or returns it, GPPComplex.cu:239, gppCudaVer.cpp:316).
 */

/* Occupied bands n1 < nvband: achtemp_re/_im[3], asxtemp[3], acht_n1_loc[nvband].
 * Replaces d_till_nvband, GPPComplex.h:306 / GPPComplex.cu:406-418. */
gpp_status d_till_nvband(int nvband, int ngpown, int ncouls, int *inv_igp_index, double *wx_array,
                         gpp_complex *wtilde_array, gpp_complex *aqsmtemp, gpp_complex *aqsntemp,
                         gpp_complex *I_eps_array, gpp_complex *asxtemp, double *vcoul,
                         gpp_complex *acht_n1_loc, double *achtemp_re, double *achtemp_im,
                         const double occ);

/* achstemp over all bands (reduce_achstemp, gppComplex.cpp:66-101).
 * Declared but never defined by the reference, GPPComplex.h:308. */
gpp_status d_reduce_achstemp(int number_bands, int *inv_igp_index, int ncouls,
                             gpp_complex *aqsmtemp, gpp_complex *aqsntemp,
                             gpp_complex *I_eps_array, double *achstemp_re, double *achstemp_im,
                             int ngpown, double *vcoul);

/* Unoccupied bands nvband <= n1 < number_bands: achtemp_re/_im[3],
 * acht_n1_loc[number_bands].  Replaces d_noFlagOCCSolver, GPPComplex.h:310 /
 * GPPComplex.cu:422-433. */
gpp_status d_noFlagOCCSolver(int number_bands, int nvband, int ngpown, int ncouls,
                             int *inv_igp_index, double *wx_array, gpp_complex *wtilde_array,
                             gpp_complex *aqsmtemp, gpp_complex *aqsntemp,
                             gpp_complex *I_eps_array, double *vcoul, double *achtemp_re,
                             double *achtemp_im, gpp_complex *acht_n1_loc);

/*
 * d_noFlagOCCSolver without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (harnesses
 * that forbid hipMalloc, tight serving loops, code that must not grow the
 * device heap).  Same parameters, meaning, accumulation and status as
 * d_noFlagOCCSolver; its scratch (index flag, results and partial sums, about
 * 16 B x number_bands x ceil(ngpown/4) x 3 plus small tables) lives in a
 * 16 MiB device array of the library, one per device, held until the process
 * ends -- enough for number_bands x ngpown up to about 512 x 1638 or 1024 x
 * 1000.  Larger problems are refused with GPP_ERR_INVALID_ARGUMENT.  Like
 * d_noFlagOCCSolver it synchronises once to validate inv_igp_index and then
 * enqueues on the null stream.  Calls on one device share the array (each
 * overwrites what the last left there): the library serialises its own calls,
 * so calls that overlap on one device -- from several host threads, or before
 * the previous call's work has run -- keep their order on the null stream.
 */
gpp_status d_noFlagOCCSolver_noalloc(int number_bands, int nvband, int ngpown, int ncouls,
                                     int *inv_igp_index, double *wx_array, gpp_complex *wtilde_array,
                                     gpp_complex *aqsmtemp, gpp_complex *aqsntemp,
                                     gpp_complex *I_eps_array, double *vcoul, double *achtemp_re,
                                     double *achtemp_im, gpp_complex *acht_n1_loc);

#ifdef __cplusplus
}
#endif

#endif /* GPP_H */
