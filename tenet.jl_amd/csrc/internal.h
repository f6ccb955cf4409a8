// Internal declarations shared by the host runtime (tenet_hip.cpp) and the
// HIP kernel files.  Nothing here crosses the C-ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace tenet {

// ---- error plumbing ------------------------------------------------------
void set_error(const std::string& msg);
#define TENET_HIP_CHECK(expr)                                                                  \
  do {                                                                                         \
    hipError_t _e = (expr);                                                                    \
    if (_e != hipSuccess) {                                                                    \
      ::tenet::set_error(std::string(#expr) + " failed: " + hipGetErrorString(_e));            \
      return TENET_ERR_HIP;                                                                    \
    }                                                                                          \
  } while (0)

// ---- K1: gather-GEMM over offset tables ------------------------------------
// C[rowC[m]+colC[n]+bC[z]] = alpha * sum_k A[rowA[m]+kA[k]+bA[z]] * B[kB[k]+colB[n]+bB[z]] + beta*C
struct GemmTables {
  const int64_t* rowA;
  const int64_t* kA;
  const int64_t* bA;
  const int64_t* kB;
  const int64_t* colB;
  const int64_t* bB;
  const int64_t* rowC;
  const int64_t* colC;
  const int64_t* bC;
};

struct GemmLaunch {
  const void* A;
  const void* B;
  void* C;
  GemmTables t;
  int64_t M, N, K, batch;
  int conjA, conjB;
  int a_mfast;  // rowA has unit stride: load A tiles with consecutive lanes along m
  int b_nfast;  // colB has unit stride: load B tiles with consecutive lanes along n
  double alpha_re, alpha_im, beta_re, beta_im;
};

// prec: 0 native (fp64 for c128), 1 fp32, 2 split-bf16, 3 split-fp16, 4 / 5 split-int8 (7 / 5 slices)
int launch_gemm_c128(const GemmLaunch& g, int prec, hipStream_t s, void* ws, size_t ws_bytes);
int launch_gemm_c64(const GemmLaunch& g, int prec, hipStream_t s, void* ws, size_t ws_bytes);
size_t gemm_workspace_bytes(const GemmLaunch& g, int dtype, int prec);
// int8 split precision (k_split8.hip)
int launch_split8_gemm(const GemmLaunch& g, int dtype, int prec, hipStream_t s, void* ws, size_t ws_bytes);
size_t split8_workspace_bytes(const GemmLaunch& g, int dtype, int prec);

// Grouped GEMM: problems and work units in device memory, one persistent
// launch.  Operand pointers come separately (3 per problem: A, B, C) so one
// plan serves many launches.
enum { GROUPED_TILE_64 = 0, GROUPED_TILE_32 = 1, GROUPED_TILE_16 = 2, GROUPED_SMALLK = 3 };
// GROUPED_SMALLK units: SMALLK_ELEMS consecutive output elements each, one
// thread per element (contractions with K <= 8 and an output edge below 16:
// Hadamard scalings, gate applications, strided copies, where an MFMA tile
// would be mostly padding)
constexpr int64_t SMALLK_ELEMS = 4096;
struct GroupedProblem {
  GemmTables t;
  int64_t M, N, K, batch;
  int32_t a_mfast, b_nfast;
  int32_t conjA, conjB;     // conjugate operand A / B
  int32_t shape;            // GROUPED_TILE_* / GROUPED_SMALLK
  int32_t kchunks;          // K chunks per tile (1: no split)
  int32_t order;            // GROUPED_SMALLK: fastest index class (0: m, 1: n, 2: batch), C's unit-stride one
  int32_t pad0;
  int64_t tiles_m, tiles_n;
  int64_t kc;               // chunk length (multiple of 16)
  int64_t part_base;        // first partial element (complex) of this problem
  int64_t ctr_base;         // first tile arrival counter of this problem
};
struct GroupedUnit {
  int32_t prob;
  int32_t chunk;
  int64_t tile;  // z * tiles_m * tiles_n + tn * tiles_m + tm
};
struct GroupedLaunch {
  const GroupedProblem* probs;
  const GroupedUnit* units;
  int64_t nunits;
  const uint64_t* ptrs;
  void* partials;
  uint32_t* counters;  // [0] queue head, [64 + i] tile arrival counters
  int64_t grid;
  int dtype, prec;
};
int launch_grouped_gemm(const GroupedLaunch& g, hipStream_t s);
int grouped_occupancy(int dtype);  // resident workgroups per CU of the grouped kernel

// ---- K2: Householder QR ----------------------------------------------------
// In-place QR of column-major W (m x n, ld = m) for `batch` matrices laid out
// back to back (stride m*n); tau (k per matrix).  Then Q (m x k) / R (k x n).
int launch_qr_factor(double* W, double* tau, int64_t m, int64_t n, int64_t batch, hipStream_t s);
// flags: TENET_QR_Q_ADJ writes Q^H (k x m, ldq); TENET_QR_R_ADJ writes R^H (n x k, ldr)
int launch_qr_extract(const double* W, const double* tau, int64_t m, int64_t n, int64_t batch, void* Q,
                      int64_t ldq, void* R, int64_t ldr, int dtype, int flags, hipStream_t s);

// single-workgroup register-resident QR (m, n <= 64), reading A in place;
// flags: TENET_QR_IN_ADJ, TENET_QR_Q_ADJ, TENET_QR_R_ADJ.  Rout may be null.
// abs (optional): R absorbed into a neighbour in the same launch, side 0:
// Bout (k x nb) = R B (B n x nb), side 1: Bout (nb x k) = B R^H (B nb x n).
struct QrAbsorbArgs {
  const void* B;
  void* Bout;
  int64_t ldb, ldbo, nb;
  int side;
};
bool qr_small_fits(int64_t m, int64_t n);
int launch_qr_small(const void* A, int64_t lda, int64_t m, int64_t n, void* Qout, int64_t ldq, void* Rout,
                    int64_t ldr, int dtype, int flags, const QrAbsorbArgs* abs, hipStream_t s,
                    hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
// Batched blocked QR (k_qr.hip): problem i factors its m x n complex128 W in
// place (R on and above the diagonal, reflectors below), keeping tau (min(m,
// n) complex) and the panels' T factors (qr_batched_t_bytes); Yp is the
// trailing update's partial products (qr_batched_yp_bytes with the batch's
// largest n).  Rows at most 512.
struct QrBatchProb {
  double* W;
  double* tau;
  double* T;
  double* Yp;
  double* TB;  // T of the current 64-column block reflector (qr_batched_tb_bytes)
  int32_t m, n;
};
size_t qr_batched_yp_bytes(int64_t m, int64_t max_n);
size_t qr_batched_tb_bytes();
size_t qr_batched_t_bytes(int64_t m, int64_t n);
int launch_qr_batched(const QrBatchProb* probs, int64_t nprob, int max_m, int max_n, int max_k, hipStream_t s);
bool qr_blocked_fits(int64_t m);
size_t qr_blocked_ws_bytes(int64_t m, int64_t n);
int launch_qr_blocked(double* W, int64_t m, int64_t n, void* Qout, int64_t ldq, void* Rout, int64_t ldr, int dtype,
                      int flags, hipStream_t s);

// ---- K3: one-sided block Jacobi SVD -----------------------------------------
struct SvdProblem {
  double* X;      // nr x ncp work matrix (complex, column-major)
  double* V;      // ncp x ncp accumulated rotations
  double* Gpart;  // npairs * q partial Gram blocks (2b x 2b complex)
  double* sigma;  // ncp
  int32_t* perm;  // ncp
  int32_t* flag;  // [0] rotated in this sweep, [1] converged, [2] sweeps; flag + 4: the sweep's max squared relative off-diagonal (double)
  double* fro2;   // ||A||_F^2 (absolute rotation floor eps^2 * ||A||_F^2)
  int32_t nr, ncols, ncp, b, nblk, q;
  int32_t m, n, transposed;
  int32_t accum;  // apply the rotations to V (0: V is never needed, U == NULL with X = A^H)
  int32_t s_mode; // S output: 0 double, 1 complex128 (sigma, 0), 2 complex64
  const void* A;  // input (m x n, lda)
  int64_t lda;
  void* U;
  int64_t ldu;
  void* Vh;
  int64_t ldvh;
  double* S;
  int64_t maxdim;
  double threshold;
  int64_t* k_dev;
  double* disc_dev;
  double tol;
  // QR preconditioning (grouped problems with m >= n and no U): W = A[:, cperm]
  // (m x n complex128, columns of A by descending norm) = Q R in place, the
  // Jacobi runs on X = R^H and V^H's columns are written through cperm
  double* W;
  int32_t* cperm;
  int32_t pre;
  int32_t pad1;
};
// Problems are passed either as a device array (grouped) or inline (one).
// units (grouped rounds): when set, a one-dimensional grid whose workgroup i
// takes (problem, pair) = units[i] instead of (blockIdx.y, blockIdx.x)
struct SvdBatch {
  const SvdProblem* arr;
  SvdProblem one;
  const int2* units;
};
int launch_svd_init(const SvdBatch& b, int64_t nprob, int64_t max_elems, int max_ncp, int dtype, hipStream_t s);
// ev0 / ev1 (optional): stamped at the kernel's start and end by the dispatch
int launch_svd_round(const SvdBatch& b, int64_t nprob, int max_items, int round, int phase, int mode,
                     hipStream_t s, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
int launch_svd_sweep_end(const SvdBatch& b, int64_t nprob, hipStream_t s);
int launch_svd_finalize(const SvdBatch& b, int64_t nprob, int max_ncp, int64_t max_elems, int dtype,
                        hipStream_t s);
// bit 0 of *status <- 1 for every problem not flagged converged
int launch_svd_status(const SvdBatch& b, int64_t nprob, int32_t* status, hipStream_t s);
// single matrix: T = A (adj 0) or A^H (adj 1), rows x cols; squared column
// norms into nrm, their descending order into perm, Tp = T[:, perm] (c128)
int launch_pre1_order_pack(const void* A, int64_t lda, int rows, int cols, int adj, int dtype, double* nrm,
                           int32_t* perm, double* Tp, hipStream_t s);
// complex128 src (rows x cols, ld lds) through the column order perm: mode 0
// dst[:, perm[c]] = src[:, c]; mode 1 dst[perm[c], r] = conj(src[r, c])
int launch_perm_scatter(const double* src, int rows, int cols, int64_t lds, double* dst, int64_t ldd,
                        const int32_t* perm, int mode, hipStream_t s);
// column norms of A, their descending order into cperm and W = A[:, cperm]
// for the problems with pre set (the others return at once)
int launch_svd_pre_pack(const SvdBatch& b, int64_t nprob, int max_n, int64_t max_elems, int dtype, hipStream_t s);

// ---- misc elementwise / K4 ----------------------------------------------------
int launch_fill_zero(void* dst, size_t bytes, hipStream_t s);
// Gate update epilogue: Aout = U[:, :k] * f(S), Bout = compact(f(S) * Vh[:k, :]).
int launch_absorb(const void* U, int64_t m, const void* Vh, int64_t ldvh, int64_t n, const double* S,
                  const int64_t* k_dev, int absorb, void* Aout, void* Bout, int64_t kmax, int dtype,
                  hipStream_t s);
// scale by 1/max|x| and accumulate log(max) into acc[2]; optionally mantissa out
int launch_rescale_log(void* x, int64_t count, double* acc, int dtype, hipStream_t s);
// site-parallel amplitudes (k_misc.hip): device tables as for
// launch_amplitudes, plus chi_max / d bounds and workspace rows (two
// 2 chi_max double rows per request, one log-scale per request).
struct AmpSiteArgs {
  const int64_t* n;
  const uint64_t* sites;
  const uint64_t* chis;
  const uint64_t* bits;
  int64_t d, chi_max, n_max;
  double* vbuf;
  double* logs;
  double* out;
};
// one site position of every request (a no-op for requests with fewer sites);
// e0 / e1: optional event pair stamped by the dispatch
int launch_amp_site(const AmpSiteArgs& a, int64_t nreq, int64_t site, int dtype, hipStream_t s, hipEvent_t e0,
                    hipEvent_t e1);
int launch_amplitudes(int64_t nreq, const int64_t* n_dev, const uint64_t* site_ptrs, const uint64_t* chis,
                      const uint64_t* bits, int64_t d, double* out, int dtype, int max_chi, hipStream_t s);
int launch_write_scalar_from(const void* env, double* out, int dtype, hipStream_t s);
// <b|a> in one workgroup when n <= 96, every bond <= 32 and d chi^2 <= 2048; the
// bond sizes and site pointers travel as the kernel's argument (3 KiB)
constexpr int OS_NMAX = 96;
struct OverlapTab {
  int32_t n, d;
  int64_t chi_a[OS_NMAX + 1], chi_b[OS_NMAX + 1];
  uint64_t a[OS_NMAX], b[OS_NMAX];
};
bool overlap_small_fits(int64_t n, int64_t d, const int64_t* chi_a, const int64_t* chi_b);
int launch_overlap_small(const OverlapTab& tab, double* out, int dtype, hipStream_t s);
int launch_trace_diag(const void* env, int64_t chi, double* out, int dtype, hipStream_t s);
int launch_convert(const void* src, int src_dtype, void* dst, int dst_dtype, int64_t count, hipStream_t s);
int launch_pack_c128(const void* src, int dtype, int64_t m, int64_t n, int64_t lda, int adj, void* dst,
                     hipStream_t s);
int launch_apply_pauli(const void* A, int64_t chil, int64_t chir, char which, void* out, int dtype, hipStream_t s);
int launch_set_identity(void* out, int64_t chi, int dtype, hipStream_t s);
int launch_gaussian(void* out, int dtype, int64_t count, uint64_t seed, uint64_t offset, hipStream_t s);
int launch_scale(void* x, int dtype, int64_t count, double ar, double ai, hipStream_t s);
int launch_sample_select(const void* T, int dtype, int d, int chir, int64_t nsamp, int n, int site, uint64_t seed,
                         int32_t* bits, void* Vn, hipStream_t s);
int launch_nrm2(const void* x, int dtype, int64_t count, double* part, double* out, hipStream_t s);
int launch_sigma_scales(const double* S, int64_t k, void* sq, void* isq, int dtype, hipStream_t s);
// Two-site gates on a batch of bonds (passed by value: no table upload, so
// the launch can be captured): per problem theta~[l, p, q, r] =
// sum_{x, y} G[p, q, x, y] theta[l, x, y, r] and, with lam set,
// X[l, p, q, r] = lam[l] theta~[l, p, q, r] (column-major, d x d gate).
struct GateProb {
  const void* theta;
  const void* gate;
  const void* lam;  // may be null (no X)
  void* theta_out;  // must not overlap theta
  void* x_out;      // may equal theta (each element is read before it is written, by the same thread)
  int64_t chil, chir;
};
constexpr int GATE_BATCH = 32;
struct GateBatch {
  GateProb p[GATE_BATCH];
  int32_t count;
  int32_t d;
};
int launch_gate_apply(const GateBatch& b, int64_t max_elems, int dtype, hipStream_t s);

}  // namespace tenet
