/*
 * csk.h -- C ABI of libcsk.so, the MI355X (gfx950) Groth16 prover core that
 * stands behind cusnarks' pycusnarks binding for the GPU multi-exponentiation
 * and polynomial path.
 *
 * Conventions (identical to cusnarks, src/cuda/types.h:150-233):
 *   - a field element is 8 x uint32 little-endian limbs;
 *   - curve coordinates are Montgomery form mod q (R = 2^256);
 *   - G1 affine point = [X | Y] (16 words), G2 = [X.c0 | X.c1 | Y.c0 | Y.c1]
 *     (32 words), Fq2 = Fq[u]/(u^2 + 1);
 *   - affine infinity is (0, 1*R); (0, 0) inputs are also read as infinity;
 *   - multi-exponentiation scalars are normal (non-Montgomery) form, any
 *     256-bit value (reduced mod r internally: every point handed to the
 *     multiexp must lie in the order-r subgroup, which all proving-key points
 *     and every BN254 G1 point do);
 *   - pidx 0 = base field q, pidx 1 = scalar field r.
 *
 * Device-pointer entry points ("_dev" absent, pointers named d_*) take memory
 * allocated on the context's device and a hipStream_t passed as void*
 * (NULL = the null stream); they enqueue work and return without
 * synchronising.  "_host" entry points copy host buffers in and out and
 * return after the result is on the host.  Every entry point returns
 * CSK_OK (0) or a negative CSK_ERR_* code; csk_last_error() describes the
 * last failure of the calling thread.  Nothing here calls exit().
 *
 * A context is not thread-safe: use one per host thread (or serialise).
 */
#ifndef CSK_H
#define CSK_H
#include <stdint.h>
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

#define CSK_OK 0
#define CSK_ERR_ARG (-1)
#define CSK_ERR_HIP (-2)
#define CSK_ERR_NOMEM (-3)
#define CSK_ERR_NODEV (-4)

typedef struct csk_ctx csk_ctx;

/* ---------------------------------------------------------------- context */
/* Replaces CUSnarks::CUSnarks / init_h (cusnarks src/cuda/cusnarks_kernel.cu:107-127,
 * src/cuda/init.cpp:44-49): binds a device and owns its scratch memory. */
int csk_ctx_create(int device, csk_ctx **out);
void csk_ctx_destroy(csk_ctx *ctx);
const char *csk_last_error(void);
int csk_device_count(int *count);

/* Per-stage device time of the last multiexp on this context, recorded with
 * HIP events on its launch stream while timing is on: ms[0] digits + sort,
 * ms[1] bucket accumulation, ms[2] fix-up + reductions + final combine,
 * ms[3] total.  Missing values are -1. */
int csk_ctx_set_timing(csk_ctx *ctx, int on);
int csk_last_stage_times(csk_ctx *ctx, float *ms, int n);
/* Plan of the last single multiexp: c, W, buckets/window, chunk, reduce
 * segment, segments/window, n, entry bound (first n values written). */
int csk_last_msm_plan(csk_ctx *ctx, uint32_t *out, int n);

/* Plain device memory helpers for callers without their own allocator. */
int csk_malloc(csk_ctx *ctx, void **d_ptr, uint64_t bytes);
int csk_free(csk_ctx *ctx, void *d_ptr);
int csk_memcpy_h2d(csk_ctx *ctx, void *d_dst, const void *src, uint64_t bytes, void *stream);
int csk_memcpy_d2h(csk_ctx *ctx, void *dst, const void *d_src, uint64_t bytes, void *stream);
int csk_memcpy_d2d(csk_ctx *ctx, void *d_dst, const void *d_src, uint64_t bytes, void *stream);
int csk_stream_sync(csk_ctx *ctx, void *stream);
/* Page-locked host memory for host buffers that are copied to the device
 * often, such as a prover's witness: copies from it run at the link's full
 * rate (a 2^24-entry witness, 512 MB, from ordinary host memory moves at
 * about 27 GB/s).  No counterpart in the reference, whose witness is read
 * into ordinary host memory; the caller fills the buffer in place. */
int csk_host_alloc(csk_ctx *ctx, void **h_ptr, uint64_t bytes);
int csk_host_free(csk_ctx *ctx, void *h_ptr);

/* Streams (default priority, non-blocking) and events for callers that
 * overlap stages without a framework of their own.  event_query returns 1
 * when the event has completed, 0 while pending, < 0 on error. */
int csk_stream_create(csk_ctx *ctx, void **stream);
int csk_stream_destroy(csk_ctx *ctx, void *stream);
int csk_stream_wait_event(csk_ctx *ctx, void *stream, void *event);
int csk_event_create(csk_ctx *ctx, void **event);
int csk_event_destroy(csk_ctx *ctx, void *event);
int csk_event_record(csk_ctx *ctx, void *event, void *stream);
int csk_event_query(csk_ctx *ctx, void *event);
int csk_event_sync(csk_ctx *ctx, void *event);
int csk_event_elapsed(csk_ctx *ctx, void *start, void *end, float *ms);

/* ------------------------------------------------------ multi-exponentiation */
/* Tuning knobs; zero fields pick the built-in defaults for the size. */
typedef struct {
  uint32_t window_bits; /* Pippenger window c (2..20) */
  uint32_t chunk;       /* sorted (bucket, point) entries per accumulation thread */
  uint32_t reduce_seg;  /* buckets per bucket-reduction thread */
  uint32_t reserved;
} csk_msm_cfg;

/* out = sum_i k_i * P_i as an affine Montgomery point (16 words G1, 32 words G2).
 * Replaces the GPU multiexp ECBN128/EC2BN128.kernelLaunch with
 * kernel_idx CB_EC_JAC_MAD_SHFL (cusnarks src/cython/_cusnarks_kernel.pyx:72-192,
 * src/cuda/cusnarks_kernel.cu:402-573, kernel src/cuda/ecbn128_device.cu:508-565)
 * followed by ec_jac2aff_h, and the CPU multiexp ec_jacreduce_h
 * (src/cython/_cusnarks_kernel.pyx:992-1053, src/cuda/ec.cpp:1536-1704). */
int csk_msm_g1(csk_ctx *ctx, const uint32_t *d_scalars, const uint32_t *d_points,
               uint64_t n, uint32_t *d_out, const csk_msm_cfg *cfg, void *stream);
int csk_msm_g2(csk_ctx *ctx, const uint32_t *d_scalars, const uint32_t *d_points,
               uint64_t n, uint32_t *d_out, const csk_msm_cfg *cfg, void *stream);
int csk_msm_g1_host(csk_ctx *ctx, const uint32_t *scalars, const uint32_t *points,
                    uint64_t n, uint32_t *out, const csk_msm_cfg *cfg);
int csk_msm_g2_host(csk_ctx *ctx, const uint32_t *scalars, const uint32_t *points,
                    uint64_t n, uint32_t *out, const csk_msm_cfg *cfg);

/* csk_msm_g1 without any device allocation, for callers that cannot call
 * hipMalloc (code under HIP graph capture, tight serving loops, harnesses
 * that forbid allocation).  It has no counterpart in the reference.  The
 * pipeline's workspace is a fixed 48 MiB device array inside the library, one
 * copy per GPU, held until the process ends.  n <= 2^15 points; the window
 * width is the tuned one narrowed until the workspace fits (the sum is the
 * same for every width); an explicit cfg->window_bits whose workspace does not
 * fit, or a larger n, returns CSK_ERR_ARG.  All calls on one GPU share the
 * array: calls that overlap on one GPU (other streams, other contexts, other
 * threads) race and must be serialised by the caller. */
int csk_msm_g1_noalloc(csk_ctx *ctx, const uint32_t *d_scalars, const uint32_t *d_points,
                       uint64_t n, uint32_t *d_out, const csk_msm_cfg *cfg, void *stream);

/* Per-size-class settings used when a call passes no cfg (or zero fields):
 * table[lg] applies to sizes n with floor(log2 n) == lg, for lg < count <= 64;
 * all-zero entries keep the built-in choice.  Written by the autotuner
 * (cusnarks_amd/tune.py).  Replaces the reference's fixed launch geometry
 * per multiexp batch (get_shfl_blockD, src/python/cuda_wrapper.py:1570-1587). */
int csk_msm_set_tuning(csk_ctx *ctx, int ec2, const csk_msm_cfg *table, uint32_t count);

/* ------------------------------------------------- batched multi-exponentiation
 * Many independent multiexps of uneven sizes on one curve (ec2 = 0 G1, 1 G2)
 * in one pipeline: members with the same window width form a group whose
 * stages each run as one launch over all members; groups run concurrently on
 * the batch's own streams, and the whole sequence can be captured once into a
 * HIP graph and replayed.  Replaces the reference's chunked dispatch of the
 * prover's multiexps over GPUs and streams (findECPointsDispatch / init_EC_P,
 * src/python/groth_prover.py:1548-1688; buildDispatchTable,
 * src/python/cuda_wrapper.py:1643-1677), whose results it reproduces exactly.
 *
 * create: plans the groups for sizes n[0..count) (window_bits[i] optional,
 *   0 = tuned/built-in) and allocates the device memory the batch needs.
 *   With flags & CSK_BATCH_CTX_WORKSPACE the large workspace is instead the
 *   context's batch workspace, grown on demand and shared by every such batch
 *   of the context (runs of those batches must not overlap; use one context
 *   per concurrent lane).  Only the small descriptor tables are owned.
 * run: out[i] = sum_j d_scalars[i][j] * d_points[i][j] (affine Montgomery,
 *   16 / 32 words), enqueued on stream.  Pointers may change between runs.
 *   use_graph != 0 replays a graph captured on the first such run (captured
 *   again if the shared workspace moved).  Runs of one batch must not overlap.
 * bytes: the workspace bytes such a batch needs (no allocation).
 * info: out = {groups, device bytes, then per group: c, members, chunk,
 *   entry bound} (first n values). */
#define CSK_BATCH_CTX_WORKSPACE 1u
typedef struct csk_msm_batch csk_msm_batch;
int csk_msm_batch_create(csk_ctx *ctx, int ec2, uint32_t count, const uint64_t *n,
                         const uint32_t *window_bits, uint32_t flags, csk_msm_batch **out);
int csk_msm_batch_run(csk_msm_batch *b, const uint32_t *const *d_scalars,
                      const uint32_t *const *d_points, uint32_t *const *d_out,
                      int use_graph, void *stream);
int csk_msm_batch_bytes(csk_ctx *ctx, int ec2, uint32_t count, const uint64_t *n,
                        const uint32_t *window_bits, uint64_t *bytes);
int csk_msm_batch_info(const csk_msm_batch *b, uint64_t *out, int n);
void csk_msm_batch_destroy(csk_msm_batch *b);

/* ------------------------------------------------------- fixed-base tables
 * For points that stay fixed across many multiexps (proving keys), a table
 * of window-shifted copies Q[w][i] = 2^(c w) P_i (w < floor(254/c) + 1) lets
 * every call accumulate all windows' digits into one set of 2^(c-1) buckets:
 * one bucket reduction per call and no doublings between windows.  Replaces
 * the reference's precomputed EC tables for fixed points (setup write_tables,
 * cusnarks src/python/groth_setup.py:802-945; ec_jacreduce_batch_precomputed_h,
 * src/cuda/ec.cpp:1837-1979; scmulecjac_precomputed_kernel,
 * src/cuda/ecbn128_device.cu:267-294), whose results it reproduces exactly.
 *
 * create: builds the table of n affine Montgomery points (ec2 = 0 G1, 1 G2)
 *   for window width window_bits (0 = a default for n), allocating
 *   floor(254/c) + 1 copies of the points (64 / 128 B each) plus one bit per
 *   point; returns after the table is built.
 * run: out = sum_i d_scalars[i] * P_i (affine Montgomery, 16 / 32 words),
 *   enqueued on stream, in the context's workspace (runs on one context must
 *   not overlap).
 * run_batch: the runs for count <= 64 scalar vectors (d_scalars[p] ->
 *   d_out[p], each result equal to run's), as one pipeline: one sort, one
 *   accumulation launch and one window combination for all of them, in the
 *   workspace of ctx (null: the table's context), which may differ from the
 *   table's so that runs on several streams do not share a workspace.  It has
 *   the role of the reference's prover looping its precomputed-table
 *   multiexp over the proofs it serves (src/python/groth_prover.py:792-891).
 * run_multi: count <= 64 runs over possibly different tables of one curve
 *   and window width (tables[p], d_scalars[p] -> d_out[p]) as one pipeline,
 *   in ctx's workspace: the multiexps of one proof over a key's tables
 *   (pysnarkP_CPU's multiexp calls, src/python/groth_prover.py:575-790).
 * bytes: the table bytes for (ec2, n, window_bits) without allocating.
 * info: out = {n, window bits, windows, table bytes, workspace bytes}. */
typedef struct csk_msm_table csk_msm_table;
int csk_msm_table_create(csk_ctx *ctx, int ec2, const uint32_t *d_points, uint64_t n, uint32_t window_bits,
                         csk_msm_table **out, void *stream);
int csk_msm_table_run(csk_msm_table *t, const uint32_t *d_scalars, uint32_t *d_out, void *stream);
int csk_msm_table_run_batch(csk_ctx *ctx, csk_msm_table *t, uint32_t count, const uint32_t *const *d_scalars,
                            uint32_t *const *d_out, void *stream);
int csk_msm_table_run_multi(csk_ctx *ctx, uint32_t count, csk_msm_table *const *tables,
                            const uint32_t *const *d_scalars, uint32_t *const *d_out, void *stream);
int csk_msm_table_bytes(int ec2, uint64_t n, uint32_t window_bits, uint64_t *bytes);
int csk_msm_table_info(const csk_msm_table *t, uint64_t *out, int n);
void csk_msm_table_destroy(csk_msm_table *t);

/* -------------------------------------------------------------- field ops */
/* Elementwise ops over n elements of field pidx; replaces the U256 kernels
 * (cusnarks src/cuda/u256_device.cu:48-610: CB_U256_MULM, ADDM, SUBM, ...) and
 * the host helpers montmultN_h / to_montgomeryN_h / from_montgomeryN_h
 * (src/cython/_cusnarks_kernel.pyx:317-331, 1087-1103). */
#define CSK_OP_MULM 0      /* out = a*b*R^-1 */
#define CSK_OP_ADDM 1      /* out = a+b */
#define CSK_OP_SUBM 2      /* out = a-b */
#define CSK_OP_TOMONT 3    /* out = a*R */
#define CSK_OP_FROMMONT 4  /* out = a*R^-1 */
#define CSK_OP_INVM 5      /* out = (aR)^-1 R (Montgomery inverse, 0 -> 0) */
#define CSK_OP_MULM_EXT 6  /* Fq2 product; n counts Fq2 elements (16 words) */
int csk_field_op(csk_ctx *ctx, int op, uint32_t pidx, const uint32_t *d_a,
                 const uint32_t *d_b, uint32_t *d_out, uint64_t n, void *stream);

/* Jacobian -> affine for n points (replaces ec_jac2aff_h / ec2_jac2aff_h,
 * src/cuda/ec.cpp:963-1062); strip_last as in the reference. */
int csk_ec_jac2aff(csk_ctx *ctx, uint32_t ec2, const uint32_t *d_jac, uint64_t n,
                   uint32_t *d_out, uint32_t strip_last, void *stream);

/* Point validation for keys read from files: counts the n affine Montgomery
 * points (ec2 = 0 G1, 1 G2) that have a coordinate >= q, are not on the curve
 * (y^2 = x^3 + 3, twist y^2 = x^3 + 3/(9+u)) or, with subgroup != 0, fail
 * [r]P == O (needed for G2; every BN254 G1 point is in the subgroup).  Both
 * infinity encodings pass.  out[0] = bad points, out[1] = lowest bad index
 * (2^64-1 if none); returns after the counts are on the host.  Replaces the
 * setup's ec_isoncurve_h checks (cusnarks src/python/groth_setup.py:744-770,
 * ec_isoncurve_h src/cuda/ec.cpp:1064) for the keys the prover loads. */
int csk_ec_check(csk_ctx *ctx, uint32_t ec2, const uint32_t *d_points, uint64_t n, uint32_t subgroup,
                 uint64_t *out, void *stream);

/* Fixed-base k_i * G (G1 generator (1,2) or the standard G2 generator) as affine
 * Montgomery points: builds synthetic proving keys and test inputs (the role of
 * ec_sc1mul_cuda in setup, cusnarks src/python/groth_setup.py:596-732). */
int csk_fixed_base_g1(csk_ctx *ctx, const uint32_t *d_scalars, uint64_t n,
                      uint32_t *d_out, void *stream);
int csk_fixed_base_g2(csk_ctx *ctx, const uint32_t *d_scalars, uint64_t n,
                      uint32_t *d_out, void *stream);

/* ------------------------------------------------------------ polynomials */
/* Fr NTT of 2^log2n Montgomery elements, natural order in and out (d_in may
 * equal d_out; log2n <= 27).  inverse = 0: X[k] = sum a[i] w^(ik), w the
 * 2^log2n-th root 5^((r-1)/2^log2n) (replaces ntt_h, cusnarks
 * src/cuda/ntt.cpp:453-463, and the ZCUPoly FFT kernels).  inverse = 1: the
 * inverse transform scaled by N^-1, Montgomery output if out_fmt = 1, normal
 * if 0 (replaces intt_h, src/cuda/ntt.cpp:583-603). */
int csk_ntt(csk_ctx *ctx, const uint32_t *d_in, uint32_t *d_out, uint32_t log2n,
            int inverse, int out_fmt, void *stream);

/* H step: d_out (2m elements, normal form) = coefficients of A(x) B(x), where
 * A, B are the degree < m interpolants of the m Montgomery evaluations d_pA,
 * d_pB on the m-th roots of unity (2 <= m <= 2^26).  Replaces
 * ntt_interpolandmul_h (src/cython/_cusnarks_kernel.pyx:479-504,
 * src/cuda/ntt.cpp:672-1070); the prover keeps d_out[m .. 2m-2]. */
int csk_interpolandmul(csk_ctx *ctx, const uint32_t *d_pA, const uint32_t *d_pB,
                       uint32_t log2m, uint32_t *d_out, void *stream);

/* H steps of `count` proofs of one circuit size at once: member t computes
 * d_out[t] from d_pA[t], d_pB[t] exactly as csk_interpolandmul does, with
 * each transform pass launched once for up to 32 members, so the transforms
 * of small circuits fill the GPU together (stage-level batching of the H
 * step across proofs in flight; the reference computes one H per proof,
 * ntt_interpolandmul_server_h, src/cuda/ntt.cpp:672-765).  Workspace: 4m
 * elements per member of a group of 32. */
int csk_interpolandmul_batch(csk_ctx *ctx, uint32_t count, const uint32_t *const *d_pA,
                             const uint32_t *const *d_pB, uint32_t log2m, uint32_t *const *d_out, void *stream);

/* QAP evaluation: d_out[row] = sum_e val[e] * toMont(wit[col[e]]) over the
 * row's CSR entries (val Montgomery, wit normal form, out Montgomery).
 * Replaces mpoly_eval_h (src/cython/_cusnarks_kernel.pyx:452-477,
 * src/cuda/mpoly.cpp:125-168), whose per-wire layout the host converts to CSR. */
int csk_mpoly_eval_csr(csk_ctx *ctx, const uint64_t *d_rowptr, const uint32_t *d_col,
                       const uint32_t *d_val, const uint32_t *d_wit, uint32_t nrows,
                       uint32_t *d_out, void *stream);

/* ------------------------------------------------------------ measurement */
/* Integer-VALU throughput probe (no counterpart in the reference): kind 0
 * measures v_mad_u64_u32 per second, kind 1 Fq Montgomery products per second
 * (kind 2: the same with a compiler-scheduled CIOS product, for comparison),
 * kind 3 bytes per second (read + written) of a plain 1 GiB device copy;
 * ms receives the best kernel time. */
int csk_valu_probe(csk_ctx *ctx, uint32_t kind, double *ops_per_s, double *ms);

#ifdef __cplusplus
}
#endif
#endif
