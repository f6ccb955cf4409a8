/*
 * tenet_hip.h -- C-ABI of libtenet_hip.so, the MI355X (gfx950) engine for
 * Tenet.jl's contraction / QR-SVD truncation / transfer-matrix hot path.
 *
 * Conventions
 *  - Every entry point returns an int status (tenet_status); on failure the
 *    thread-local message is available from tenet_last_error().
 *  - Device buffers are plain pointers obtained from tenet_malloc (or any
 *    device allocation of the same device).  Complex data are interleaved
 *    (re, im) pairs: TENET_C128 = Julia ComplexF64, TENET_C64 = ComplexF32.
 *  - Matrices and tensors are column-major (Julia layout) unless a stride
 *    descriptor says otherwise.
 *  - A NULL tenet_stream means the context's own stream.
 *  - No entry point falls back to the CPU: without a usable gfx950 device
 *    tenet_create fails with TENET_ERR_NODEVICE.
 *
 * The reference interfaces replaced are Muscle.jl's Tensor kernels as called
 * from Tenet.jl v0.10.3 (Muscle is not vendored; the call sites are cited).
 */
#ifndef TENET_HIP_H
#define TENET_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TENET_MAX_RANK 8

typedef enum {
  TENET_OK = 0,
  TENET_ERR_HIP = 1,         /* a HIP runtime call failed */
  TENET_ERR_SHAPE = 2,       /* labels / extents / strides inconsistent */
  TENET_ERR_OOM = 3,         /* device allocation failed or budget exceeded */
  TENET_ERR_NOCONV = 4,      /* Jacobi SVD hit its sweep limit */
  TENET_ERR_UNSUPPORTED = 5, /* valid request this build does not implement */
  TENET_ERR_ARG = 6,         /* bad argument (null pointer, bad enum, ...) */
  TENET_ERR_NODEVICE = 7     /* no gfx950 device visible */
} tenet_status;

typedef enum { TENET_C128 = 0, TENET_C64 = 1 } tenet_dtype;

/* Arithmetic of a contraction.  NATIVE = the dtype's own precision
 * (FP64 for C128, FP32 for C64).  SPLIT_BF16 / SPLIT_FP16 = operands split
 * into 8-bit slices (bf16 / fp16 MFMA), exact FP32 accumulation in short K
 * blocks, FP64 summation across blocks and slices.  SPLIT_I8 = 7-bit slices
 * on int8 MFMA with exact int32 accumulation (seven slices for complex128
 * results: FP64 grade, 2^-49; four for complex64); SPLIT_I8_FAST = five
 * slices for complex128 results (2^-35 relative to the row / column
 * maxima).  The result is written in the tensor dtype. */
typedef enum {
  TENET_PREC_NATIVE = 0,
  TENET_PREC_FP32 = 1,
  TENET_PREC_SPLIT_BF16 = 2,
  TENET_PREC_SPLIT_FP16 = 3,
  TENET_PREC_SPLIT_I8 = 4,
  TENET_PREC_SPLIT_I8_FAST = 5
} tenet_prec;

typedef enum { TENET_ABSORB_NONE = 0, TENET_ABSORB_EQUAL = 1, TENET_ABSORB_LEFT = 2, TENET_ABSORB_RIGHT = 3 } tenet_absorb;

typedef struct tenet_ctx_s* tenet_ctx;
typedef struct tenet_plan_s* tenet_plan;
typedef struct tenet_graph_s* tenet_graph;
typedef void* tenet_stream; /* hipStream_t */

/* Labelled dense tensor: element (i_0..i_{r-1}) lives at sum_d i_d*stride[d]
 * (in elements).  Labels are arbitrary int64 tags (Tenet's Index tags). */
typedef struct {
  int32_t rank;
  int32_t reserved;
  int64_t extent[TENET_MAX_RANK];
  int64_t stride[TENET_MAX_RANK];
  int64_t label[TENET_MAX_RANK];
} tenet_tdesc;

/* ---- library / context ------------------------------------------------- */
const char* tenet_version(void);
const char* tenet_last_error(void);
int tenet_device_count(int* count);
int tenet_create(int device, tenet_ctx* ctx);
int tenet_destroy(tenet_ctx ctx);
int tenet_stream_create(tenet_ctx ctx, tenet_stream* stream);
int tenet_stream_destroy(tenet_ctx ctx, tenet_stream stream);
int tenet_stream_sync(tenet_ctx ctx, tenet_stream stream);
int tenet_device_sync(tenet_ctx ctx);

/* ---- memory (stream-ordered caching pool owned by the context) ---------- */
int tenet_malloc(tenet_ctx ctx, size_t bytes, void** dptr);
int tenet_free(tenet_ctx ctx, void* dptr);
int tenet_memcpy_h2d(tenet_ctx ctx, void* dst, const void* src, size_t bytes, tenet_stream stream);
int tenet_memcpy_d2h(tenet_ctx ctx, void* dst, const void* src, size_t bytes, tenet_stream stream);
int tenet_memcpy_d2d(tenet_ctx ctx, void* dst, const void* src, size_t bytes, tenet_stream stream);
int tenet_memset_zero(tenet_ctx ctx, void* dst, size_t bytes, tenet_stream stream); /* fill kernel */
int tenet_mem_info(tenet_ctx ctx, size_t* free_bytes, size_t* total_bytes, size_t* pool_bytes);
/* Stream-ordered free: the block returns to the pool once the work queued on
 * `stream` (NULL: the context's stream) before this call has completed, so it
 * may be called while kernels on that stream still use the block. */
int tenet_free_async(tenet_ctx ctx, void* dptr, tenet_stream stream);
/* Two-phase workspace: the bytes one call of `op` on an m x n matrix
 * (TENET_OP_SIMPLE_UPDATE: the bond matrix, m = chil*d1, n = d2*chir) takes
 * from the per-stream workspace, and a call that grows that workspace now
 * (so that a later graph capture, which cannot grow it, finds it large
 * enough). */
typedef enum { TENET_OP_QR = 1, TENET_OP_SVD = 2, TENET_OP_SIMPLE_UPDATE = 3 } tenet_op;
int tenet_workspace_size(tenet_ctx ctx, int op, int dtype, int64_t m, int64_t n, size_t* bytes);
int tenet_workspace_reserve(tenet_ctx ctx, tenet_stream stream, size_t bytes);
/* Library log: level 0 off, 1 every error message as it is set, 2 also
 * resource events (workspace growth); to `path` (appended) or stderr when
 * path is NULL or empty.  Process-wide. */
int tenet_set_log(int level, const char* path);

/* ---- asynchronous request loops (no reference counterpart: the serving
 * loop's plumbing) --------------------------------------------------------
 * Pinned host buffers (hipHostMalloc) for copies that return before they
 * complete; events to poll completion and order streams.  Pool memory freed
 * with tenet_free is reused at once, so free a buffer only after the work
 * using it is known complete. */
typedef void* tenet_event; /* hipEvent_t */
int tenet_host_alloc(tenet_ctx ctx, size_t bytes, void** hptr);
int tenet_host_free(tenet_ctx ctx, void* hptr);
/* to_host = 1: device -> pinned host; 0: pinned host -> device.  Asynchronous. */
int tenet_memcpy_async(tenet_ctx ctx, void* dst, const void* src, size_t bytes, int to_host, tenet_stream stream);
int tenet_event_create(tenet_ctx ctx, tenet_event* ev);
int tenet_event_destroy(tenet_ctx ctx, tenet_event ev);
int tenet_event_record(tenet_ctx ctx, tenet_event ev, tenet_stream stream);
int tenet_event_query(tenet_ctx ctx, tenet_event ev, int* done); /* done = 1 when complete */
int tenet_stream_wait_event(tenet_ctx ctx, tenet_stream stream, tenet_event ev);

/* ---- pairwise contraction (K1) ------------------------------------------
 * Replaces Muscle.binary_einsum(a, b; dims) at Tenet call sites
 * src/Operations/canonize.jl:44,47,61,100; evolve.jl:29,57,76,86,108,118;
 * overlap.jl:26,28,42,46; simple_update.jl:34; Algorithms/DMRG.jl:17,298,300.
 * Label classes: contracted = (a ∩ b) \ c; batch (Hadamard) = a ∩ b ∩ c;
 * free = labels of one operand, all of which must appear in c.
 * dims=Index[] of the reference is a c that keeps every shared label.
 * C = alpha * op(A) op(B) + beta * C, op = identity or conjugate. */
int tenet_einsum_plan(tenet_ctx ctx, int dtype, int prec, const tenet_tdesc* a, const tenet_tdesc* b,
                      const tenet_tdesc* c, tenet_plan* plan);
int tenet_einsum_exec(tenet_plan plan, tenet_stream stream, const void* A, const void* B, void* C, int conj_a,
                      int conj_b, const double alpha[2], const double beta[2]);
int tenet_plan_destroy(tenet_plan plan);
/* plan (cached by descriptor) + exec in one call */
int tenet_einsum(tenet_ctx ctx, tenet_stream stream, int dtype, int prec, const tenet_tdesc* a, const void* A,
                 const tenet_tdesc* b, const void* B, const tenet_tdesc* c, void* C, int conj_a, int conj_b);

/* ---- thin QR (K2) ---------------------------------------------------------
 * Replaces Muscle.tensor_qr_thin at src/Operations/canonize.jl:58.
 * A is m x n (column-major, lda >= m); k = min(m, n).  Q: m x k (ldq),
 * R: k x n upper triangular (ldr).  Householder reflections in FP64. */
int tenet_qr_thin(tenet_ctx ctx, tenet_stream stream, int dtype, int64_t m, int64_t n, const void* A, int64_t lda,
                  void* Q, int64_t ldq, void* R, int64_t ldr);
/* Operand options for the same factorisation, used by right-to-left canonize
 * moves (canonize.jl:142-145) so no transposed copy is made:
 * IN_ADJ: factorise A^H, where the given A is n x m (lda >= n);
 * Q_ADJ: write Q^H (k x m, ldq >= k); R_ADJ: write R^H (n x k, ldr >= n). */
typedef enum { TENET_QR_IN_ADJ = 1, TENET_QR_Q_ADJ = 2, TENET_QR_R_ADJ = 4 } tenet_qr_flags;
int tenet_qr_thin_ex(tenet_ctx ctx, tenet_stream stream, int dtype, int flags, int64_t m, int64_t n, const void* A,
                     int64_t lda, void* Q, int64_t ldq, void* R, int64_t ldr);
/* One canonize move in one launch: the QR above followed by R's contraction
 * into the neighbour tensor, replacing generic_canonize_site!'s
 * tensor_qr_thin + binary_einsum(R, B) (src/Operations/canonize.jl:58-70).
 * flags: TENET_QR_IN_ADJ, TENET_QR_Q_ADJ.  R itself is not returned.
 * side TENET_ABSORB_R_B: Bout (k x nb, ldbo >= k) = R B, B n x nb (ldb >= n);
 * side TENET_ABSORB_B_RH: Bout (nb x k, ldbo >= nb) = B R^H, B nb x n
 * (ldb >= nb).  Bout must not overlap A, Q or B, nor Q overlap B.  m, n <= 64 (the
 * single-workgroup QR); larger shapes return TENET_ERR_UNSUPPORTED (use
 * tenet_qr_thin_ex + tenet_einsum). */
typedef enum { TENET_ABSORB_R_B = 0, TENET_ABSORB_B_RH = 1 } tenet_absorb_side;
int tenet_qr_absorb(tenet_ctx ctx, tenet_stream stream, int dtype, int flags, int64_t m, int64_t n, const void* A,
                    int64_t lda, void* Q, int64_t ldq, int side, int64_t nb, const void* B, int64_t ldb, void* Bout,
                    int64_t ldbo);

/* ---- thin SVD + truncation (K3) -------------------------------------------
 * Replaces Muscle.tensor_svd_thin at canonize.jl:41,97; evolve.jl:62,92;
 * DMRG.jl:338,437 plus the truncation rule of compress.jl:46-59.
 * A = U diag(S) Vh with S descending; kmax = min(m, n).  U: m x kmax (ldu),
 * S: kmax doubles, Vh: kmax x n (ldvh).  Kept count k = min(maxdim, #S >
 * threshold) (maxdim <= 0: no cap; threshold <= 0: none); if k_dev /
 * discarded_dev are non-NULL, k (int64) and sum_{j>=k} S_j^2 (double) are
 * written to device memory.  One-sided block Jacobi, FP64 rotations. */
int tenet_svd_thin(tenet_ctx ctx, tenet_stream stream, int dtype, int64_t m, int64_t n, const void* A, int64_t lda,
                   void* U, int64_t ldu, double* S, void* Vh, int64_t ldvh, int64_t maxdim, double threshold,
                   int64_t* k_dev, double* discarded_dev);
/* Allocation-free variant of tenet_svd_thin.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, outputs and truncation as tenet_svd_thin;
 * the Jacobi workspace is a per-GPU static device array of
 * TENET_SVD_NOALLOC_SCRATCH_BYTES (held until the process ends), so no
 * hipMalloc / hipFree or allocating routine is called.  It runs the plain
 * sweeps (no QR preconditioning; every shape with min(m, n) <= 512 and
 * max(m, n) <= 1024 fits); a shape whose workspace does not fit returns
 * TENET_ERR_UNSUPPORTED.  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (two streams, two threads) race. */
#define TENET_SVD_NOALLOC_SCRATCH_BYTES ((size_t)24 << 20)
int tenet_svd_thin_noalloc(tenet_ctx ctx, tenet_stream stream, int dtype, int64_t m, int64_t n, const void* A,
                           int64_t lda, void* U, int64_t ldu, double* S, void* Vh, int64_t ldvh, int64_t maxdim,
                           double threshold, int64_t* k_dev, double* discarded_dev);
/* Complex128 matrices with 64 <= min(m, n) and max(m, n) <= 512 take a thin
 * Householder QR first and run the Jacobi on R^H (fewer sweeps, same
 * relative accuracy); on by default, 0 turns it off (tenet_svd_thin and
 * tenet_simple_update). */
int tenet_svd_set_precondition(tenet_ctx ctx, int on);

/* ---- fused two-site gate update (K1 + K3) -----------------------------------
 * Replaces Muscle.simple_update(A, ia, B, ib, bond, G, oa, ob; maxdim,
 * absorb) at src/Operations/simple_update.jl:53-64.
 * A: (chil, d1, chi), B: (chi, d2, chir), G: (d1, d2, d1, d2) holding
 * G[oa, ob, ia, ib], all column-major.  theta = A.B over chi; theta.G over
 * (ia, ib); SVD with rows (chil, oa), cols (ob, chir); truncation as in
 * tenet_svd_thin; singular values absorbed per `absorb`.
 * Outputs: Aout (chil, d1, k) and Bout (k, d2, chir), compact column-major,
 * each with capacity for kmax = min(chil*d1, d2*chir); S_out (kmax doubles,
 * may be NULL); k and the discarded weight go to k_dev / disc_dev (device,
 * may be NULL) and, if k_host is non-NULL, k is also returned to the host
 * (this synchronises the stream). */
int tenet_simple_update(tenet_ctx ctx, tenet_stream stream, int dtype, int prec, int64_t chil, int64_t d1,
                        int64_t chi, int64_t d2, int64_t chir, const void* A, const void* B, const void* G, void* Aout,
                        void* Bout, double* S_out, int64_t maxdim, double threshold, int absorb, int64_t* k_dev,
                        double* disc_dev, int64_t* k_host);

/* ---- transfer-matrix sweeps (K4) -------------------------------------------
 * overlap(a::MPS, b::MPS) = <b|a> (src/Operations/overlap.jl:36-50) with a
 * running FP64 log-scale: result = out[0:2] * exp(out[2]).  Sites are
 * (chi_l, d, chi_r) column-major; chi arrays hold n+1 bond sizes with
 * chi[0] = chi[n] = 1.  a_sites / b_sites are HOST arrays of device
 * pointers.  out is a device pointer to 3 doubles. */
int tenet_overlap(tenet_ctx ctx, tenet_stream stream, int dtype, int prec, int64_t n, int64_t d,
                  const int64_t* chi_a, const void* const* a_sites, const int64_t* chi_b,
                  const void* const* b_sites, double* out);

/* Batched amplitudes <x|psi> (overlap.jl:22-32: no conjugation of x) for
 * nreq requests, one workgroup each.  Device arrays: site_ptrs[req] points to
 * a device array of n_req device pointers; chis[req] to n_req+1 int64 bond
 * sizes; bits[req] to n_req uint8 basis indices.  out: device, 3 doubles per
 * request (re, im, log-scale). */
int tenet_amplitudes(tenet_ctx ctx, tenet_stream stream, int dtype, int64_t nreq, const int64_t* n_dev,
                     const uint64_t* site_ptrs_dev, const uint64_t* chis_dev, const uint64_t* bits_dev, int64_t d,
                     double* out);
/* Site-parallel form of tenet_amplitudes (same request tables, same
 * results within rounding): one launch per site position, each request's
 * vector-slice product spread over ceil(chi_r / 16) workgroups, so that a
 * large state (chi 512: 2 MiB per selected slice) is read by many CUs at
 * once instead of one.  n_max / chi_max bound every request's site count
 * and bond sizes (a request with more sites or a larger bond reports NaN).
 * Replaces the same call site, overlap(::ProductState, ::MPS) at
 * src/Operations/overlap.jl:22-32. */
int tenet_amplitudes_sites(tenet_ctx ctx, tenet_stream stream, int dtype, int64_t nreq, int64_t n_max,
                           int64_t chi_max, const int64_t* n_dev, const uint64_t* site_ptrs_dev,
                           const uint64_t* chis_dev, const uint64_t* bits_dev, int64_t d, double* out);

/* Pauli-string expectation <psi|P|psi> on sites [start, start+w) given the
 * left environment env_l (chi_start x chi_start, device; NULL = identity,
 * for a left-canonical prefix) and a right-canonical suffix (identity right
 * environment).  paulis: w chars from "IXYZ" (ext/TenetPythonCallExt/
 * pytket.jl:23-31 conventions).  out: device, 3 doubles (re, im, log). */
int tenet_expect_pauli(tenet_ctx ctx, tenet_stream stream, int dtype, int prec, int64_t start, int64_t w,
                       int64_t d, const int64_t* chi, const void* const* sites, const void* env_l,
                       const char* paulis, double* out);
/* Left environment E(r, r') = sum over sites [0, upto) of conj(A) A (the
 * prefix of the overlap.jl:36-50 sweep), for reuse as env_l: env_out gets E
 * divided by exp(out[2]) (chi[upto]^2 elements, device); out: device, 3
 * doubles with the log-scale in out[2].  upto = 0 gives the 1 x 1 identity. */
int tenet_left_env(tenet_ctx ctx, tenet_stream stream, int dtype, int prec, int64_t upto, int64_t d,
                   const int64_t* chi, const void* const* sites, void* env_out, double* out);
/* Every left environment of one sweep: env_outs (HOST array of upto+1
 * device pointers, NULL entries skipped) receive E(j) / exp(logs[j]) for
 * j = 0..upto (chi[j]^2 elements each); logs: device, upto+1 doubles.
 * Synchronises the stream before returning. */
int tenet_left_envs(tenet_ctx ctx, tenet_stream stream, int dtype, int prec, int64_t upto, int64_t d,
                    const int64_t* chi, const void* const* sites, void* const* env_outs, double* logs);

/* ---- grouped (persistent) launches ---------------------------------------
 * Many independent contractions (each one Muscle.binary_einsum, as at the
 * call sites of canonize.jl:44-61 / evolve.jl:29-119, batched) in one
 * persistent launch.  descs holds 3*count descriptors (a, b, c per problem).
 *
 * The plan analyses every problem once: offset tables in one device arena,
 * a tile shape per problem (64x64, 32x32 or 16x16 outputs, chosen by a cost
 * model whose per-shape costs tenet_autotune_tiles measures), long K ranges
 * cut into chunks, and a work queue of (problem, tile, chunk) units sorted
 * by decreasing cost that resident workgroups drain.  Chunk partials are
 * summed in chunk order by the workgroup finishing a tile, so results do
 * not depend on the schedule.  The plan owns its device memory. */
typedef struct tenet_gplan_s* tenet_gplan;
int tenet_grouped_plan(tenet_ctx ctx, int dtype, int prec, int64_t count, const tenet_tdesc* descs,
                       tenet_gplan* plan);
/* The same with per-problem conjugation: conj[2*i] conjugates problem i's
 * a, conj[2*i + 1] its b (binary_einsum on conj(a) / conj(b), as the
 * transfer-matrix and Vidal-gauge updates need; NULL = none). */
int tenet_grouped_plan_conj(tenet_ctx ctx, int dtype, int prec, int64_t count, const tenet_tdesc* descs,
                            const int32_t* conj, tenet_gplan* plan);
/* ptrs: HOST array of 3*count device pointers (A, B, C per problem), copied
 * before the call returns (the host array may be reused at once).  The
 * launch is asynchronous on the stream: later work on the stream (and a
 * synchronising copy to the host) sees the results; one plan's launches must
 * stay on one stream. */
int tenet_grouped_exec(tenet_gplan plan, tenet_stream stream, void* const* ptrs);
/* ptrs_dev: DEVICE array of 3*count pointers; asynchronous, graph-capturable */
int tenet_grouped_exec_dev(tenet_gplan plan, tenet_stream stream, const uint64_t* ptrs_dev);
/* shapes: problems per unit kind (64 x 64, 32 x 32, 16 x 16 MFMA tiles;
 * element-wise units for K <= 8 with an output edge below 16) */
int tenet_grouped_plan_info(tenet_gplan plan, int64_t* units, int64_t* split_tiles, int64_t shapes[4],
                            double* flops, double* bytes);
int tenet_grouped_plan_destroy(tenet_gplan plan);
/* plan + exec + destroy */
int tenet_grouped_einsum(tenet_ctx ctx, tenet_stream stream, int dtype, int prec, int64_t count,
                         const tenet_tdesc* descs, void* const* ptrs);
/* Measure the three tile shapes (64 problems of (256x1024).(1024x256)) and
 * keep their relative cost per (tile, K step) in the context; costs_out
 * (may be NULL) receives them, 64x64 = 1. */
int tenet_autotune_tiles(tenet_ctx ctx, int dtype, double costs_out[3]);
/* Many independent thin SVDs (column-major, lda = m) with per-problem
 * truncation (the tensor_svd_thin calls of canonize.jl:97 / evolve.jl:62,
 * batched).  All pointer arrays are HOST arrays of device pointers;
 * ms, ns, maxdims are host arrays; k_dev / disc_dev device arrays (count).
 * Optional: Us[i] = NULL (U not formed; a problem with m <= n then runs the
 * Jacobi on A^H and accumulates no rotations), Ss = NULL, maxdims = NULL
 * (no cap), k_dev = NULL, disc_dev = NULL.  As[i] and Vhs[i] are required.
 * Returns once the last kernels are queued on the stream (the host reads the
 * convergence flags once per Jacobi sweep, one sweep behind the launches):
 * U, S, Vh, k_dev and disc_dev are read after a later synchronisation. */
int tenet_grouped_svd(tenet_ctx ctx, tenet_stream stream, int dtype, int64_t count, const int64_t* ms,
                      const int64_t* ns, const void* const* As, void* const* Us, double* const* Ss,
                      void* const* Vhs, const int64_t* maxdims, double threshold, int64_t* k_dev, double* disc_dev);
/* The same grouped SVD bound to its operands once (shapes, problem table and
 * every Jacobi workspace in one device block), so that a run is kernel
 * launches only and can be captured (tenet_graph_*).  s_mode
 * TENET_S_DTYPE writes S as complex values of the matrices' dtype (sigma, 0),
 * so that it can scale a tensor directly (the Vidal-gauge Lambda); S then
 * holds kmax complex entries.  Eager exec: as tenet_grouped_svd.  Captured
 * exec: a fixed 30 sweeps, and the context's device status bit 0
 * (tenet_graph_status) for problems they did not finish. */
typedef struct tenet_gsvd_s* tenet_gsvd;
typedef enum { TENET_S_REAL = 0, TENET_S_DTYPE = 1 } tenet_s_mode;
int tenet_grouped_svd_plan(tenet_ctx ctx, int dtype, int64_t count, const int64_t* ms, const int64_t* ns,
                           const void* const* As, void* const* Us, void* const* Ss, void* const* Vhs,
                           const int64_t* maxdims, double threshold, int64_t* k_dev, double* disc_dev, int s_mode,
                           tenet_gsvd* plan);
int tenet_grouped_svd_exec(tenet_gsvd plan, tenet_stream stream);
int tenet_grouped_svd_destroy(tenet_gsvd plan);

/* ---- elementwise helpers ---------------------------------------------------
 * Synthetic inputs on the device: element i of dst becomes a standard
 * complex Gaussian (N(0,1) + i N(0,1))/sqrt(2) drawn from a counter-based
 * stream (splitmix64 finaliser + Box-Muller) keyed by (seed, offset + i), so
 * any slice can be regenerated on the host.  Workload generator only (no
 * reference counterpart). */
int tenet_fill_gaussian(tenet_ctx ctx, tenet_stream stream, int dtype, int64_t count, void* dst, uint64_t seed,
                        uint64_t offset);
/* x *= alpha (normalize!(tensor) at src/Operations/normalize.jl:9,26 scales
 * the centre tensor this way) */
int tenet_scale(tenet_ctx ctx, tenet_stream stream, int dtype, int64_t count, void* x, const double alpha[2]);
/* One step of Tenet's sample(psi, nsamples) (src/Operations/sample.jl:20-72)
 * at site `site` of an n-site MPS canonised to its first site.  T (d x chir
 * x nsamp, column-major) holds every sample's projected site tensor
 * T[x, r, s] = sum_l V[l, s] A[l, x, r] (tenet_einsum of the previous
 * projections V with the site).  Per sample: p_x = sum_r |T[x, r, s]|^2,
 * x* drawn from p / sum(p) with the uniform of generator element
 * s * n + site (the tenet_fill_gaussian stream's u1, seed `seed`),
 * bits[s * n + site] = x* (int32), and, unless Vn is NULL (last site),
 * Vn[r, s] = T[x*, r, s] / sqrt(p_x*).  d <= 8. */
int tenet_sample_select(tenet_ctx ctx, tenet_stream stream, int dtype, const void* T, int d, int chir, int64_t nsamp,
                        int n, int site, uint64_t seed, int32_t* bits, void* Vn);
/* sqrt_out[j] = sqrt(S[j]) and inv_sqrt_out[j] = 1 / sqrt(S[j]) (0 where
 * S[j] = 0) as complex vectors of length k: the diagonal factors of a bond
 * after simple_update's AbsorbEqually (simple_update.jl:63), with which the
 * next canonize! move across that bond is a diagonal rescaling (the thin QR
 * of U sqrt(S) is U times sqrt(S), up to diagonal phases). */
int tenet_sigma_scales(tenet_ctx ctx, tenet_stream stream, int dtype, int64_t k, const double* S, void* sqrt_out,
                       void* inv_sqrt_out);
/* Two-site gates on `count` bonds at once: the gate contraction inside
 * Muscle.simple_update (simple_update.jl:53-64, the contraction of the gate's
 * inputs with theta) batched over the disjoint bonds of a brickwork layer,
 * fused with the Vidal-gauge scaling by the left bond's Lambda.  Per
 * problem i (column-major, d = physical dimension, 2..4):
 *   theta_out[l, p, q, r] = sum_{x, y} G[p, q, x, y] theta[l, x, y, r]
 * with theta (chil x d x d x chir) and G (d x d x d x d, G[oa, ob, ia, ib]:
 * the 4x4 U4[d oa + ob, d ia + ib] of simple_update!), and, when lams[i] is
 * set, x_out[l, p, q, r] = lams[i][l] theta_out[l, p, q, r] (lams[i] a
 * vector of chil values of the dtype).  lams and x_out may be NULL, or hold
 * NULL entries (no X for that problem; both or neither).  x_out[i] may equal
 * theta[i] (computed in place); theta_out[i] must not overlap theta[i].
 * Asynchronous; the problem table travels in the launch (capturable). */
int tenet_grouped_gate(tenet_ctx ctx, tenet_stream stream, int dtype, int64_t count, int d, const int64_t* chil,
                       const int64_t* chir, const void* const* theta, const void* const* gates,
                       const void* const* lams, void* const* theta_out, void* const* x_out);
/* out (device double) = sum |x_i|^2 in FP64 (norm(tensor), norm.jl:28) */
int tenet_nrm2sq(tenet_ctx ctx, tenet_stream stream, int dtype, int64_t count, const void* x, double* out);

/* ---- HIP graph capture of a sequence of calls on one stream ---------------
 * Captured sweeps: the calls between tenet_graph_begin and tenet_graph_end on
 * `stream` (NULL: the context's stream) are recorded, not run; a launch
 * replays them (NULL: on the capture stream).  The canonize! / overlap /
 * norm sweeps (src/Operations/canonize.jl:136-145, overlap.jl:36-50) capture
 * whole; run a sequence once eagerly first, so that workspaces and plans
 * exist (neither may grow during a capture).  Calls that read results back
 * to the host (tenet_grouped_svd, tenet_simple_update with k_host, the
 * memcpy helpers) cannot be captured.  A captured tenet_svd_thin runs a
 * fixed 30 sweeps (rounds of a converged matrix return at once) and, if
 * they did not converge, sets bit 0 of the context's device status word:
 * tenet_graph_status synchronises the context stream, returns that word and
 * clears it. */
int tenet_graph_begin(tenet_ctx ctx, tenet_stream stream);
int tenet_graph_end(tenet_ctx ctx, tenet_stream stream, tenet_graph* graph);
int tenet_graph_launch(tenet_graph graph, tenet_stream stream);
int tenet_graph_status(tenet_ctx ctx, int* status);
int tenet_graph_destroy(tenet_graph graph);

/* ---- diagnostics -------------------------------------------------------- */
/* Per-kernel timing: while enabled for a kernel family, every 17th launch of
 * that family (see tenet_stats_sample_every) is bracketed by HIP events on
 * its stream and the algorithmic
 * work of every launch is accumulated (flops: complex MAC = 8; bytes:
 * operands read + written once).  tenet_stats_read synchronises and returns
 * the launch count, the timed total scaled to all launches, and the work
 * totals since the last read, then clears them. */
typedef enum {
  TENET_KSTAT_OFF = 0,
  TENET_KSTAT_SVD_ROUND = 1, /* one-sided Jacobi round (Gram, rotations, apply) */
  TENET_KSTAT_GEMM = 2,      /* contraction kernels (all precisions) */
  TENET_KSTAT_QR = 3,        /* one thin QR: pack, panels, trailing and Q updates */
  TENET_KSTAT_AMPLITUDE = 4  /* batched amplitude sweep (time only: the work
                                depends on device-side tables) */
} tenet_kstat;
int tenet_stats_enable(tenet_ctx ctx, int kernel_family);
int tenet_stats_read(tenet_ctx ctx, int64_t* launches, double* total_ms, double* flops, double* bytes);
/* Event-bracket every `every`-th launch of the tracked family (default 17;
 * 1 times every launch, for short runs of unlike launches). */
int tenet_stats_sample_every(tenet_ctx ctx, int every);
/* Average duration (ms) of the last `tenet_timer_*` bracket, measured with
 * HIP events on the given stream. */
int tenet_timer_start(tenet_ctx ctx, tenet_stream stream);
int tenet_timer_stop(tenet_ctx ctx, tenet_stream stream, float* ms);

#ifdef __cplusplus
}
#endif
#endif /* TENET_HIP_H */
