/*
 * gnt — Goldilocks-NTT big-integer multiply on AMD Instinct MI355X (gfx950).
 *
 * C ABI of libgnt.so.  Plain pointers and sizes only; limb arrays are
 * little-endian uint64 magnitudes (the reference's `&[u64]` / `Vec<u64>`).
 * Every entry point returns 0 on success or a negative GNT_E_* code and never
 * aborts; gnt_last_error() gives a thread-local message.  A context is bound
 * to one device and is safe to share between host threads.
 *
 * Reference interface (joelodom/pi, crates/bignum):
 *   ntt.rs:806        pub(crate) fn mul_mag_ntt(a: &[u64], b: &[u64]) -> Vec<u64>
 *   integer.rs:460    fn mul_mag_dispatch(a, b, ntt_max_coeffs, ntt_threshold, karatsuba_threshold)
 *   integer.rs:480    the only production call of mul_mag_ntt (guarded by ntt_fits && smaller >= ntt_threshold)
 */
#ifndef GNT_H
#define GNT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GNT_OK 0
#define GNT_E_ARG (-1)       /* bad argument (null pointer, capacity too small, unknown handle) */
#define GNT_E_TOO_LARGE (-2) /* 4*(na+nb) exceeds the transform cap: caller splits (integer.rs:479-485) */
#define GNT_E_NOMEM (-3)     /* device memory or the HBM admission cap exhausted */
#define GNT_E_DEVICE (-4)    /* HIP runtime error */
#define GNT_E_TIMEOUT (-5)   /* gnt_wait timed out; the request is still in flight */
#define GNT_E_STATE (-6)     /* context shut down, or serving loop not running */

typedef struct gnt_ctx gnt_ctx;
typedef struct gnt_buf gnt_buf; /* device-resident limb array owned by the library */

typedef struct gnt_config {
    int device;                 /* HIP device ordinal */
    size_t hbm_cap_bytes;       /* admission cap for serving/batches; 0 = 200 GB */
    const char *plan_cache_path;/* autotuner plan cache (JSON); NULL = built-in plans */
    size_t serve_split_log_n;   /* serving: products with transforms up to 2^this run on the
                                   small-request lane, larger ones on the large lane; 0 = 22 */
    size_t serve_large_inflight;/* serving: products the large lane advances at once (admitted
                                   in arrival order); 0 = 1, SIZE_MAX = no limit */
    size_t serve_device_limbs;  /* serving: requests whose result has at most this many limbs
                                   (products na+nb <= 2048, sums and differences of operands up to
                                   2047 limbs) run in the device-resident loop; 0 = 2048,
                                   GNT_SERVE_DEVICE_OFF = none */
    size_t reserved[1];
} gnt_config;
#define GNT_SERVE_DEVICE_OFF ((size_t)-1)

/* ---- lifetime ---- */
const char *gnt_version(void);
const char *gnt_last_error(void);
/* Replaces bignum::config::apply (config.rs:148) for this path: one context per device. */
int gnt_init(gnt_ctx **ctx, const gnt_config *cfg);
int gnt_shutdown(gnt_ctx *ctx);

/* ---- single products (host buffers) ----
 * gnt_mul replaces ntt::mul_mag_ntt (ntt.rs:806-865): out receives |a|*|b|
 * trimmed of most-significant zero limbs (ntt.rs:712-715, 781-783);
 * out_cap >= na+nb.  na == 0 or nb == 0 gives *out_len = 0 (ntt.rs:807-809).
 * GNT_E_TOO_LARGE when 4*(na+nb) > 2^32 coefficients, the reference's cap
 * (integer.rs:476-478); the caller splits with Karatsuba (gnt_mul_mag does).
 * Thread-safe; while the serving loop runs (gnt_serve_start), calls from any
 * thread are submitted to it, so concurrent callers are batched together. */
int gnt_mul(gnt_ctx *ctx, const uint64_t *a, size_t na, const uint64_t *b, size_t nb, uint64_t *out,
            size_t out_cap, size_t *out_len);
/* gnt_sqr replaces the reference's squares, which all go through mul_mag(a, a)
 * (float.rs:182, integer.rs:78,95, gauss_legendre.rs:130); one forward
 * transform instead of two, same product.  out_cap >= 2*na. */
int gnt_sqr(gnt_ctx *ctx, const uint64_t *a, size_t na, uint64_t *out, size_t out_cap, size_t *out_len);

/* mul_mag / mul_mag_dispatch (integer.rs:445-486) with the GPU multiply as
 * its NTT leaf: schoolbook below karatsuba_threshold, Karatsuba (recursing
 * through this dispatch) in between and above the transform cap, the GPU
 * product when it fits and min(na, nb) >= ntt_threshold.  Zero thresholds
 * select the reference defaults (8192, 32).  out_cap >= na+nb+1. */
int gnt_mul_mag(gnt_ctx *ctx, const uint64_t *a, size_t na, const uint64_t *b, size_t nb, uint64_t *out,
                size_t out_cap, size_t *out_len, size_t ntt_threshold, size_t karatsuba_threshold);
/* mul_mag_dispatch (integer.rs:460-486) with the transform cap passed in, as
 * the reference passes ntt_max_coeffs so that its tests can force the
 * Karatsuba-above-NTT branch on small inputs (integer.rs:1727-1754): the GPU
 * product is taken only when 4(na+nb) <= ntt_max_coeffs (0 or anything above
 * the library cap 2^32 means the cap).  Karatsuba's three sub-products run
 * concurrently while the smaller operand has >= 512 limbs (rayon::join in
 * integer.rs:565-593).  karatsuba_threshold 1 is refused (it never ends). */
int gnt_mul_mag_dispatch(gnt_ctx *ctx, const uint64_t *a, size_t na, const uint64_t *b, size_t nb, uint64_t *out,
                         size_t out_cap, size_t *out_len, size_t ntt_max_coeffs, size_t ntt_threshold,
                         size_t karatsuba_threshold);

/* ---- device-resident operands (no reference counterpart: the reference has
 * no device; these keep operands and results in HBM between dependent
 * products, as LimbStorage (storage.rs:33-44) keeps them in host memory) ---- */
int gnt_buf_upload(gnt_ctx *ctx, const uint64_t *limbs, size_t n, gnt_buf **out);
int gnt_buf_download(gnt_ctx *ctx, gnt_buf *buf, uint64_t *out, size_t cap, size_t *out_len);
int gnt_buf_len(gnt_ctx *ctx, gnt_buf *buf, size_t *len); /* trimmed length; waits for the producer */
int gnt_buf_release(gnt_ctx *ctx, gnt_buf *buf);
/* Product of two device-resident operands into a new buffer (a == b squares). */
int gnt_mul_buf(gnt_ctx *ctx, gnt_buf *a, gnt_buf *b, gnt_buf **out);
int gnt_sqr_buf(gnt_ctx *ctx, gnt_buf *a, gnt_buf **out);
/* Wait until all work queued on the context's stream has finished. */
int gnt_sync(gnt_ctx *ctx);

/* gnt_mul_buf's product for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  Operands d_a (na limbs) and d_b (nb limbs;
 * NULL squares d_a), the output d_out (out_cap >= na + nb limbs) and the word
 * d_len that receives the trimmed length are caller-owned device memory.  The
 * product's scratch (coefficient arrays, limb sums, carry flags) is a static
 * device array of GNT_NOALLOC_SCRATCH_BYTES per GPU, held until the process
 * ends, so products are limited to 4(na+nb) <= 2^GNT_NOALLOC_MAX_LOG_N
 * (GNT_E_TOO_LARGE above).  Stream-ordered on `stream` (a hipStream_t; NULL =
 * the context's stream); returns once the work is enqueued.  No hipMalloc,
 * hipFree or other allocating call; every strided pass steps its twiddles
 * (no twiddle tables).  All calls on one GPU share its scratch array: calls
 * whose work overlaps on the device race. */
#define GNT_NOALLOC_MAX_LOG_N 22
#define GNT_NOALLOC_SCRATCH_BYTES 83887104 /* 2 x 32 MiB coefficients + 16 MiB limb sums + 1 KiB flags */
int gnt_mul_noalloc(gnt_ctx *ctx, const uint64_t *d_a, size_t na, const uint64_t *d_b, size_t nb, uint64_t *d_out,
                    size_t out_cap, uint32_t *d_len, void *stream);

/* ---- batched products (within a binary-splitting node the reference runs its
 * four products one after another, chudnovsky.rs:411-413; independent
 * subtrees run concurrently through rayon::join, chudnovsky.rs:397-408) ---- */
typedef struct gnt_job {
    const uint64_t *a;
    size_t na;
    const uint64_t *b; /* ignored when is_sqr */
    size_t nb;
    uint64_t *out;
    size_t out_cap;
    size_t out_len;  /* written */
    int is_sqr;
    int status;      /* written: GNT_OK or an error code */
} gnt_job;
/* All jobs in as few launches as the plans allow; returns when all finished. */
int gnt_mul_batch(gnt_ctx *ctx, gnt_job *jobs, size_t n);
/* Same jobs, one product at a time (the unbatched path, for comparison). */
int gnt_mul_batch_serial(gnt_ctx *ctx, gnt_job *jobs, size_t n);

/* Batch plans: a fixed set of product shapes with device-resident operand,
 * scratch and result buffers, scheduled once (products of at most 2048 result
 * limbs in one persistent launch whose workgroups claim them and compute each
 * whole in LDS; every level of the other products grouped into one launch per
 * kernel shape) and optionally captured as a hipGraph; replay with new
 * operands of the same shapes.  No reference counterpart: they serve the
 * per-level products of binary splitting. */
typedef struct gnt_batch gnt_batch;
int gnt_batch_create(gnt_ctx *ctx, const size_t *na, const size_t *nb, const int *is_sqr, size_t n, int use_graph,
                     gnt_batch **out);
/* Copy jobs[i].a / .b (host) into the plan's operand buffers (shapes must match). */
int gnt_batch_upload(gnt_batch *b, const gnt_job *jobs);
/* Enqueue one run: mode 0 = batched (graph replay when captured), 1 = one
 * launch sequence per product.  Returns without waiting. */
int gnt_batch_run(gnt_batch *b, int mode);
/* Wait, then copy results into jobs[i].out (trimmed; jobs[i].out_len set). */
int gnt_batch_download(gnt_batch *b, gnt_job *jobs);
int gnt_batch_info(gnt_batch *b, size_t *launches_per_run, size_t *iterations, size_t *device_bytes);
int gnt_batch_destroy(gnt_batch *b);

/* ---- serving loop (continuous batching) ----
 * Requests with at most gnt_config.serve_device_limbs result limbs run in a
 * device-resident loop: a persistent kernel whose workgroups claim them from
 * a ring in pinned host memory and compute each whole in LDS (gnt_submit
 * copies host operands into the ring; no launch per request).
 * Other requests go to two lanes (transforms up to 2^gnt_config.serve_split_log_n, and larger),
 * each a scheduler thread on its own stream, advance all their admitted
 * requests one stage per iteration (one launch per kernel shape per
 * iteration) and admit waiting requests, in the order their operands are on
 * the device, between iterations while the bytes in flight of both lanes stay
 * under gnt_config.hbm_cap_bytes.  gnt_submit returns once the operands'
 * device buffers exist; an uploader thread copies host operands, the request
 * with the fewest bytes left first.  Host operands must stay valid until the
 * request completes.  An operand may instead be a gnt_buf (e.g. an earlier
 * result), which likewise must not be released before the request completes.
 * No reference counterpart: the reference has no serving mode. */
enum {
    GNT_OP_MUL = 0, /* a * b */
    GNT_OP_SQR = 1, /* a * a (b ignored) */
    GNT_OP_ADD = 2, /* a + b */
    GNT_OP_SUB = 3  /* a - b; requires a >= b (the result is a - b mod 2^(64*na) otherwise) */
};
typedef struct gnt_req {
    int op;             /* GNT_OP_*; ADD/SUB serve binary splitting's t = t_l*q_r +- p_l*t_r
                           (chudnovsky.rs:411) without a host round trip */
    const uint64_t *a;  /* host operand, or NULL when a_buf is set */
    size_t na;
    const uint64_t *b;
    size_t nb;
    gnt_buf *a_buf;     /* device operand (a finished result), or NULL */
    gnt_buf *b_buf;
} gnt_req;
int gnt_serve_start(gnt_ctx *ctx);
int gnt_serve_stop(gnt_ctx *ctx);   /* drains in-flight requests */
int gnt_submit(gnt_ctx *ctx, const gnt_req *req, uint64_t *ticket);
int gnt_poll(gnt_ctx *ctx, uint64_t ticket, int *done);
/* Wait for a request (timeout_us < 0: forever).  On success *result is a new
 * device buffer owned by the caller, its trimmed length already known (no
 * device round trip in gnt_buf_len); *latency_us = submit-to-completion. */
int gnt_wait(gnt_ctx *ctx, uint64_t ticket, int64_t timeout_us, gnt_buf **result, double *latency_us);
/* [0] iterations [1] launches [2] completed [3] peak admitted [4] peak bytes in flight [5] bytes in flight
 * [6] iterations of the small-request lane [7] of the large-request lane [8] requests run by the
 * device-resident loop [9] launches of its persistent kernel (n <= 10 values) */
int gnt_serve_stats(gnt_ctx *ctx, uint64_t *stats, int n);

/* ---- transform plans and the autotuner (no reference counterpart: the
 * reference's only transform knob is four_step_threshold, config.rs:81-89) ----
 * A plan splits a length-2^log_n transform into passes of radix 2^log_r[p]:
 * one pass up to 2^13, else 2..4 passes of radix 2^7..2^10.  Choices are
 * process-wide and apply to every context; change them while no product runs. */
/* Install a split for one length (npass 0 restores the built-in rule). */
int gnt_set_plan(int log_n, int npass, const int *log_r);
/* Load a plan cache written by gnt_tune (also done by gnt_init when
 * gnt_config.plan_cache_path is set). */
int gnt_load_plans(const char *path);
/* A strided pass whose block size is at most 2^max_log_m multiplies its
 * twiddles from a precomputed table (one per pass shape and direction, built
 * on first use, 8 * 2^log_m bytes); larger blocks step them.  -1 disables
 * tables; the default is 22.  *previous (if not null) receives the old
 * setting.  Process-wide, like the plans. */
int gnt_set_pass_tables(int max_log_m, int *previous);
/* Copies between caller (pageable) host memory and the device of at least
 * min_bytes go through pinned staging: chunks of chunk_bytes, copied on the
 * host by `threads` threads (the caller's included) while the DMA engine
 * moves the previous chunk.  threads = 0 sends every copy straight to the
 * runtime.  Defaults: 4 threads, 8 MiB chunks, 4 MiB minimum.  Process-wide;
 * no reference counterpart (the reference's multiply never leaves the host). */
int gnt_set_staging(int threads, size_t chunk_bytes, size_t min_bytes);
/* Time every valid split of each length in [min_log_n, max_log_n] (4..28) on
 * the device (median of reps products each, a multiply and a square), check
 * every candidate's product (fingerprints mod 2^64-1 and 2^61-1 and the low
 * limb; GNT_E_DEVICE if one is wrong), install the split whose multiply plus
 * square is fastest, and write the choices and both timings to cache_path
 * (JSON) when it is not NULL.  Tile shapes are fixed at 2^13 elements (two
 * 64 KiB tiles per CU; 2^12-element tiles at four per CU measured slower), so
 * a split is the whole plan. */
int gnt_tune(gnt_ctx *ctx, int min_log_n, int max_log_n, int reps, const char *cache_path);

/* ---- introspection ---- */
/* Transform plan the library uses for an na x nb product. */
int gnt_plan_info(size_t na, size_t nb, int *log_n, int *npass, int *log_r /* [4] */);
/* Per-phase device time (ms) accumulated while timing is on:
 * [0] forward passes before the last, [1] last forward pass + pointwise + first inverse pass,
 * [2] remaining inverse passes, [3] carry, [4] whole products. */
int gnt_set_timing(gnt_ctx *ctx, int on);
int gnt_get_timing(gnt_ctx *ctx, double *ms, int n, uint64_t *products);

#ifdef __cplusplus
}
#endif
#endif
