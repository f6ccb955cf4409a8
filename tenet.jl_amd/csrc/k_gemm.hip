// K1: pairwise tensor contraction as a gather-GEMM on CDNA4 MFMA.
//
// A binary contraction, once matricised, is a (batched) GEMM
//   C[m, n, z] = sum_k A[m, k, z] * B[k, n, z].
// The permutation that matricisation needs is never materialised: the host
// planner turns every label class into a table of element offsets
// (rowA[m], kA[k], bA[z], ...), and the tile loaders below gather straight
// from the caller's strided tensors into LDS ("permute-into-LDS").  Complex
// products run as four real MFMA products (no 3M trick: it cancels badly at
// the 1e-8 accuracy this path is held to).
//
// Tiles: 64x64 output per 256-thread workgroup, four waves each owning a
// 32x32 quadrant made of 2x2 16x16 MFMA blocks, K step 16.
//   FP64: v_mfma_f64_16x16x4_f64      FP32: v_mfma_f32_16x16x4_f32
// MFMA 16x16x4 operand layout (lane l): A[i = l%16][k = l/16],
// B[k = l/16][j = l%16]; result register r of lane l holds D[i][j = l%16]
// with i = 4*(l/16) + r for FP32 and i = 4*r + l/16 for FP64.
#include "../../include/tenet_hip.h"
#include "internal.h"

namespace tenet {

typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

template <typename R>
struct Acc4;
template <>
struct Acc4<double> {
  using T = d4v;
};
template <>
struct Acc4<float> {
  using T = f4v;
};

__device__ __forceinline__ d4v mfma16x16x4(double a, double b, d4v c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4v mfma16x16x4(float a, float b, f4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

constexpr int BM = 64, BN = 64, BK = 16, PADK = 4;

template <typename R>
struct GemmSmem {
  R ar[BK][BM + PADK];
  R ai[BK][BM + PADK];
  R br[BK][BN + PADK];
  R bi[BK][BN + PADK];
};

struct TileView {
  const void* A;
  const void* B;
  void* C;
  GemmTables t;
  int64_t M, N, K;
  int a_mfast, b_nfast, conjA, conjB;
  double alpha_re, alpha_im, beta_re, beta_im;
};

// One 64x64 output tile (tm, tn) of batch entry z.
template <typename S2, typename R>
__device__ __forceinline__ void gemm_tile(const TileView& p, int64_t tm, int64_t tn, int64_t z, GemmSmem<R>& sm) {
  using V4 = typename Acc4<R>::T;
  const S2* __restrict__ A = static_cast<const S2*>(p.A);
  const S2* __restrict__ B = static_cast<const S2*>(p.B);
  S2* __restrict__ C = static_cast<S2*>(p.C);
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int64_t m0 = tm * BM, n0 = tn * BN;
  const int64_t M = p.M, N = p.N, K = p.K;
  const int64_t offAz = p.t.bA[z], offBz = p.t.bB[z], offCz = p.t.bC[z];

  // Per-thread loader coordinates.  The dimension a thread keeps fixed over
  // the K loop has its table offset cached in registers.
  int la_m[4], la_k[4], lb_n[4], lb_k[4];
  int64_t la_fixed[4], lb_fixed[4];
  bool la_ok[4], lb_ok[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (p.a_mfast) {
      la_m[j] = tid & 63;
      la_k[j] = (tid >> 6) + 4 * j;
    } else {
      la_m[j] = (tid >> 4) + 16 * j;
      la_k[j] = tid & 15;
    }
    const int64_t gm = m0 + la_m[j];
    la_ok[j] = gm < M;
    la_fixed[j] = la_ok[j] ? p.t.rowA[gm] + offAz : 0;
    if (p.b_nfast) {
      lb_n[j] = tid & 63;
      lb_k[j] = (tid >> 6) + 4 * j;
    } else {
      lb_n[j] = (tid >> 4) + 16 * j;
      lb_k[j] = tid & 15;
    }
    const int64_t gn = n0 + lb_n[j];
    lb_ok[j] = gn < N;
    lb_fixed[j] = lb_ok[j] ? p.t.colB[gn] + offBz : 0;
  }

  R ra_r[4], ra_i[4], rb_r[4], rb_i[4];
  auto gload = [&](int64_t k0) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t gka = k0 + la_k[j];
      if (la_ok[j] && gka < K) {
        const S2 v = A[la_fixed[j] + p.t.kA[gka]];
        ra_r[j] = (R)v.x;
        ra_i[j] = p.conjA ? -(R)v.y : (R)v.y;
      } else {
        ra_r[j] = 0;
        ra_i[j] = 0;
      }
      const int64_t gkb = k0 + lb_k[j];
      if (lb_ok[j] && gkb < K) {
        const S2 v = B[lb_fixed[j] + p.t.kB[gkb]];
        rb_r[j] = (R)v.x;
        rb_i[j] = p.conjB ? -(R)v.y : (R)v.y;
      } else {
        rb_r[j] = 0;
        rb_i[j] = 0;
      }
    }
  };
  auto sstore = [&]() {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      sm.ar[la_k[j]][la_m[j]] = ra_r[j];
      sm.ai[la_k[j]][la_m[j]] = ra_i[j];
      sm.br[lb_k[j]][lb_n[j]] = rb_r[j];
      sm.bi[lb_k[j]][lb_n[j]] = rb_i[j];
    }
  };

  V4 accR[2][2], accI[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      accR[i][j] = V4{0, 0, 0, 0};
      accI[i][j] = V4{0, 0, 0, 0};
    }

  gload(0);
  __syncthreads();  // previous tile's readers are done with the LDS buffers
  sstore();
  __syncthreads();
  for (int64_t k0 = 0; k0 < K; k0 += BK) {
    const bool more = (k0 + BK) < K;
    if (more) gload(k0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK / 4; ++kk) {
      const int kr = kk * 4 + (lane >> 4);
      R xr[2], xi[2], yr[2], yi[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        xr[i] = sm.ar[kr][wm * 32 + i * 16 + (lane & 15)];
        xi[i] = sm.ai[kr][wm * 32 + i * 16 + (lane & 15)];
        yr[i] = sm.br[kr][wn * 32 + i * 16 + (lane & 15)];
        yi[i] = sm.bi[kr][wn * 32 + i * 16 + (lane & 15)];
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          accR[i][j] = mfma16x16x4(xr[i], yr[j], accR[i][j]);
          accR[i][j] = mfma16x16x4(-xi[i], yi[j], accR[i][j]);
          accI[i][j] = mfma16x16x4(xr[i], yi[j], accI[i][j]);
          accI[i][j] = mfma16x16x4(xi[i], yr[j], accI[i][j]);
        }
    }
    __syncthreads();
    if (more) {
      sstore();
      __syncthreads();
    }
  }

  // Epilogue: C = alpha * acc + beta * C, scattered through the C tables.
  const double ar = p.alpha_re, ai = p.alpha_im, br = p.beta_re, bi = p.beta_im;
  const bool use_beta = (br != 0.0 || bi != 0.0);
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int64_t col = n0 + wn * 32 + j * 16 + (lane & 15);
    if (col >= N) continue;
    const int64_t ccol = p.t.colC[col] + offCz;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        // D layout: FP32 16x16x4 keeps rows 4*(l/16)+r in lane l; FP64
        // 16x16x4 interleaves them as rows 4*r + l/16.
        const int rr = (sizeof(R) == 8) ? (r * 4 + (lane >> 4)) : ((lane >> 4) * 4 + r);
        const int64_t row = m0 + wm * 32 + i * 16 + rr;
        if (row >= M) continue;
        const int64_t off = p.t.rowC[row] + ccol;
        const double cr = (double)accR[i][j][r], ci = (double)accI[i][j][r];
        double outr = ar * cr - ai * ci;
        double outi = ar * ci + ai * cr;
        if (use_beta) {
          const S2 old = C[off];
          outr += br * (double)old.x - bi * (double)old.y;
          outi += br * (double)old.y + bi * (double)old.x;
        }
        S2 o;
        o.x = outr;
        o.y = outi;
        C[off] = o;
      }
    }
  }
}

// Map a physical workgroup id to a logical tile so that the tiles one XCD
// receives (ids congruent mod 8 are dispatched to the same XCD) are
// neighbours in the tile grid and share A/B panels in that XCD's L2.
__device__ __forceinline__ int64_t xcd_remap(int64_t pid, int64_t total) {
  constexpr int64_t NXCD = 8;
  if (total % NXCD != 0) return pid;
  return (pid % NXCD) * (total / NXCD) + pid / NXCD;
}

template <typename S2, typename R>
__global__ __launch_bounds__(256) void gemm_kernel(TileView p, int64_t tiles_m, int64_t tiles_n, int64_t batch) {
  __shared__ GemmSmem<R> sm;
  const int64_t per_batch = tiles_m * tiles_n;
  const int64_t total = per_batch * batch;
  const int64_t lid = xcd_remap((int64_t)blockIdx.x, total);
  const int64_t z = lid / per_batch;
  const int64_t t = lid - z * per_batch;
  gemm_tile<S2, R>(p, t % tiles_m, t / tiles_m, z, sm);
}

template <typename S2, typename R>
static int launch_gemm_t(const GemmLaunch& g, hipStream_t s) {
  TileView p;
  p.A = g.A;
  p.B = g.B;
  p.C = g.C;
  p.t = g.t;
  p.M = g.M;
  p.N = g.N;
  p.K = g.K;
  p.a_mfast = g.a_mfast;
  p.b_nfast = g.b_nfast;
  p.conjA = g.conjA;
  p.conjB = g.conjB;
  p.alpha_re = g.alpha_re;
  p.alpha_im = g.alpha_im;
  p.beta_re = g.beta_re;
  p.beta_im = g.beta_im;
  const int64_t tm = (g.M + BM - 1) / BM, tn = (g.N + BN - 1) / BN;
  const int64_t total = tm * tn * g.batch;
  if (total <= 0) return TENET_OK;
  if (total > 0x7fffffffLL) {
    set_error("contraction too large for one launch");
    return TENET_ERR_UNSUPPORTED;
  }
  hipLaunchKernelGGL((gemm_kernel<S2, R>), dim3((unsigned)total), dim3(256), 0, s, p, tm, tn, g.batch);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

int launch_split_gemm(const GemmLaunch& g, int dtype, int prec, hipStream_t s, void* ws, size_t ws_bytes);

int launch_gemm_c128(const GemmLaunch& g, int prec, hipStream_t s, void* ws, size_t ws_bytes) {
  if (g.M == 0 || g.N == 0 || g.batch == 0) return TENET_OK;
  switch (prec) {
    case TENET_PREC_NATIVE:
      return launch_gemm_t<double2, double>(g, s);
    case TENET_PREC_FP32:
      return launch_gemm_t<double2, float>(g, s);
    case TENET_PREC_SPLIT_BF16:
    case TENET_PREC_SPLIT_FP16:
      return launch_split_gemm(g, TENET_C128, prec, s, ws, ws_bytes);
  }
  set_error("unknown precision");
  return TENET_ERR_ARG;
}

int launch_gemm_c64(const GemmLaunch& g, int prec, hipStream_t s, void* ws, size_t ws_bytes) {
  if (g.M == 0 || g.N == 0 || g.batch == 0) return TENET_OK;
  switch (prec) {
    case TENET_PREC_NATIVE:
    case TENET_PREC_FP32:
      return launch_gemm_t<float2, float>(g, s);
    case TENET_PREC_SPLIT_BF16:
    case TENET_PREC_SPLIT_FP16:
      return launch_split_gemm(g, TENET_C64, prec, s, ws, ws_bytes);
  }
  set_error("unknown precision");
  return TENET_ERR_ARG;
}

// ---- grouped persistent launch ------------------------------------------------
template <typename S2, typename R>
__global__ __launch_bounds__(256) void grouped_gemm_kernel(const GroupedProblem* __restrict__ probs, int64_t nprob,
                                                           int64_t total_tiles) {
  __shared__ GemmSmem<R> sm;
  for (int64_t tile = blockIdx.x; tile < total_tiles; tile += gridDim.x) {
    // binary search the problem owning this tile
    int64_t lo = 0, hi = nprob - 1;
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) >> 1;
      if (probs[mid].tile_begin <= tile)
        lo = mid;
      else
        hi = mid - 1;
    }
    const GroupedProblem& q = probs[lo];
    TileView p;
    p.A = q.A;
    p.B = q.B;
    p.C = q.C;
    p.t = q.t;
    p.M = q.M;
    p.N = q.N;
    p.K = q.K;
    p.a_mfast = q.a_mfast;
    p.b_nfast = q.b_nfast;
    p.conjA = 0;
    p.conjB = 0;
    p.alpha_re = 1.0;
    p.alpha_im = 0.0;
    p.beta_re = 0.0;
    p.beta_im = 0.0;
    const int64_t local = tile - q.tile_begin;
    const int64_t per_batch = q.tiles_m * q.tiles_n;
    const int64_t z = local / per_batch;
    const int64_t t = local - z * per_batch;
    gemm_tile<S2, R>(p, t % q.tiles_m, t / q.tiles_m, z, sm);
  }
}

int launch_grouped_gemm(const GroupedProblem* probs_dev, int64_t nprob, int64_t total_tiles, int dtype,
                        hipStream_t s) {
  if (total_tiles <= 0) return TENET_OK;
  // persistent: a few workgroups per CU, each looping over tiles
  int dev = 0, ncu = 256;
  TENET_HIP_CHECK(hipGetDevice(&dev));
  TENET_HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const int64_t grid = std::min<int64_t>(total_tiles, (int64_t)ncu * 4);
  if (dtype == TENET_C128)
    hipLaunchKernelGGL((grouped_gemm_kernel<double2, double>), dim3((unsigned)grid), dim3(256), 0, s, probs_dev,
                       nprob, total_tiles);
  else
    hipLaunchKernelGGL((grouped_gemm_kernel<float2, float>), dim3((unsigned)grid), dim3(256), 0, s, probs_dev,
                       nprob, total_tiles);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

}  // namespace tenet
