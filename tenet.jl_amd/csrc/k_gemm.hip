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

constexpr int BK = 16, PADK = 4;

// Tile variants.  A workgroup of four waves computes one BM x BN output tile;
// the waves form a WGM x WGN grid over the tile and, when WGK > 1, also split
// each K step of 16 (k sub-steps of 4 taken round robin) with an LDS
// reduction at the end.  Each wave owns TM x TN MFMA blocks of 16 x 16.
template <int BM_, int BN_, int WGM_, int WGN_, int WGK_>
struct TileCfg {
  static constexpr int BM = BM_, BN = BN_, WGM = WGM_, WGN = WGN_, WGK = WGK_;
  static constexpr int TM = BM / (16 * WGM), TN = BN / (16 * WGN);
  static constexpr int NA = BM * BK / 256, NB = BN * BK / 256;  // loads per thread
  static_assert(WGM * WGN * WGK == 4, "four waves");
  static_assert(TM >= 1 && TN >= 1 && NA >= 1 && NB >= 1, "tile too small");
};
using Tile64 = TileCfg<64, 64, 2, 2, 1>;
using Tile32 = TileCfg<32, 32, 2, 2, 1>;
using Tile16 = TileCfg<16, 16, 1, 1, 4>;

template <typename R>
struct GemmSmem {
  R ar[BK][64 + PADK];
  R ai[BK][64 + PADK];
  R br[BK][64 + PADK];
  R bi[BK][64 + PADK];
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

template <typename R, typename Cfg>
struct TileAcc {
  using V4 = typename Acc4<R>::T;
  V4 r[Cfg::TM][Cfg::TN], i[Cfg::TM][Cfg::TN];
};

// Accumulate the (tm, tn) tile of batch entry z over k in [k_lo, k_hi).
// On return every wave holds its TM x TN blocks; for WGK > 1 the sum over
// the K-splitting waves has been folded into wave 0 (the others hold zeros).
// CONJ = false: the operands are never conjugated (p.conjA / p.conjB are
// not read), so the loads carry no per-element sign select.
template <typename S2, typename R, typename Cfg, bool CONJ = true>
__device__ __forceinline__ void gemm_accumulate(const TileView& p, int64_t tm, int64_t tn, int64_t z, int64_t k_lo,
                                                int64_t k_hi, GemmSmem<R>& sm, TileAcc<R, Cfg>& acc) {
  const bool conjA = CONJ && p.conjA, conjB = CONJ && p.conjB;
  constexpr int BM = Cfg::BM, BN = Cfg::BN, NA = Cfg::NA, NB = Cfg::NB;
  const S2* __restrict__ A = static_cast<const S2*>(p.A);
  const S2* __restrict__ B = static_cast<const S2*>(p.B);
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wk = wave % Cfg::WGK;
  const int wmn = wave / Cfg::WGK;
  const int wm = wmn % Cfg::WGM, wn = wmn / Cfg::WGM;
  const int64_t m0 = tm * BM, n0 = tn * BN;
  const int64_t M = p.M, N = p.N;
  const int64_t offAz = p.t.bA[z], offBz = p.t.bB[z];

  // Per-thread loader coordinates.  The dimension a thread keeps fixed over
  // the K loop has its table offset cached in registers.
  int la_m[NA], la_k[NA], lb_n[NB], lb_k[NB];
  int64_t la_fixed[NA], lb_fixed[NB];
  bool la_ok[NA], lb_ok[NB];
#pragma unroll
  for (int j = 0; j < NA; ++j) {
    if (p.a_mfast) {
      la_m[j] = tid % BM;
      la_k[j] = tid / BM + (256 / BM) * j;
    } else {
      la_m[j] = (tid >> 4) + 16 * j;
      la_k[j] = tid & 15;
    }
    const int64_t gm = m0 + la_m[j];
    la_ok[j] = gm < M;
    la_fixed[j] = la_ok[j] ? p.t.rowA[gm] + offAz : 0;
  }
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    if (p.b_nfast) {
      lb_n[j] = tid % BN;
      lb_k[j] = tid / BN + (256 / BN) * j;
    } else {
      lb_n[j] = (tid >> 4) + 16 * j;
      lb_k[j] = tid & 15;
    }
    const int64_t gn = n0 + lb_n[j];
    lb_ok[j] = gn < N;
    lb_fixed[j] = lb_ok[j] ? p.t.colB[gn] + offBz : 0;
  }

  R ra_r[NA], ra_i[NA], rb_r[NB], rb_i[NB];
  auto gload = [&](int64_t k0) {
#pragma unroll
    for (int j = 0; j < NA; ++j) {
      const int64_t gka = k0 + la_k[j];
      if (la_ok[j] && gka < k_hi) {
        const S2 v = A[la_fixed[j] + p.t.kA[gka]];
        ra_r[j] = (R)v.x;
        ra_i[j] = conjA ? -(R)v.y : (R)v.y;
      } else {
        ra_r[j] = 0;
        ra_i[j] = 0;
      }
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const int64_t gkb = k0 + lb_k[j];
      if (lb_ok[j] && gkb < k_hi) {
        const S2 v = B[lb_fixed[j] + p.t.kB[gkb]];
        rb_r[j] = (R)v.x;
        rb_i[j] = conjB ? -(R)v.y : (R)v.y;
      } else {
        rb_r[j] = 0;
        rb_i[j] = 0;
      }
    }
  };
  auto sstore = [&]() {
#pragma unroll
    for (int j = 0; j < NA; ++j) {
      sm.ar[la_k[j]][la_m[j]] = ra_r[j];
      sm.ai[la_k[j]][la_m[j]] = ra_i[j];
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      sm.br[lb_k[j]][lb_n[j]] = rb_r[j];
      sm.bi[lb_k[j]][lb_n[j]] = rb_i[j];
    }
  };

#pragma unroll
  for (int i = 0; i < Cfg::TM; ++i)
#pragma unroll
    for (int j = 0; j < Cfg::TN; ++j) {
      acc.r[i][j] = typename TileAcc<R, Cfg>::V4{0, 0, 0, 0};
      acc.i[i][j] = typename TileAcc<R, Cfg>::V4{0, 0, 0, 0};
    }

  gload(k_lo);
  __syncthreads();  // previous tile's readers are done with the LDS buffers
  sstore();
  __syncthreads();
  for (int64_t k0 = k_lo; k0 < k_hi; k0 += BK) {
    const bool more = (k0 + BK) < k_hi;
    if (more) gload(k0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK / 4; ++kk) {
      if (kk % Cfg::WGK != wk) continue;
      const int kr = kk * 4 + (lane >> 4);
      R xr[Cfg::TM], xi[Cfg::TM], yr[Cfg::TN], yi[Cfg::TN];
#pragma unroll
      for (int i = 0; i < Cfg::TM; ++i) {
        xr[i] = sm.ar[kr][(wm * Cfg::TM + i) * 16 + (lane & 15)];
        xi[i] = sm.ai[kr][(wm * Cfg::TM + i) * 16 + (lane & 15)];
      }
#pragma unroll
      for (int j = 0; j < Cfg::TN; ++j) {
        yr[j] = sm.br[kr][(wn * Cfg::TN + j) * 16 + (lane & 15)];
        yi[j] = sm.bi[kr][(wn * Cfg::TN + j) * 16 + (lane & 15)];
      }
#pragma unroll
      for (int i = 0; i < Cfg::TM; ++i)
#pragma unroll
        for (int j = 0; j < Cfg::TN; ++j) {
          acc.r[i][j] = mfma16x16x4(xr[i], yr[j], acc.r[i][j]);
          acc.r[i][j] = mfma16x16x4(-xi[i], yi[j], acc.r[i][j]);
          acc.i[i][j] = mfma16x16x4(xr[i], yi[j], acc.i[i][j]);
          acc.i[i][j] = mfma16x16x4(xi[i], yr[j], acc.i[i][j]);
        }
    }
    __syncthreads();
    if (more) {
      sstore();
      __syncthreads();
    }
  }

  if constexpr (Cfg::WGK > 1) {
    // fold the K-splitting waves into wave 0 in a fixed order (deterministic);
    // the LDS operand buffers are free after the loop's final barrier
    static_assert(Cfg::TM == 1 && Cfg::TN == 1, "K-split tiles hold one block per wave");
    R* red = &sm.ar[0][0];  // (WGK - 1) * 8 * 64 values <= the 4 operand planes
    if (wave > 0) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        red[((wave - 1) * 8 + r) * 64 + lane] = acc.r[0][0][r];
        red[((wave - 1) * 8 + 4 + r) * 64 + lane] = acc.i[0][0][r];
        acc.r[0][0][r] = 0;
        acc.i[0][0][r] = 0;
      }
    }
    __syncthreads();
    if (wave == 0) {
      for (int w = 1; w < Cfg::WGK; ++w)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          acc.r[0][0][r] += red[((w - 1) * 8 + r) * 64 + lane];
          acc.i[0][0][r] += red[((w - 1) * 8 + 4 + r) * 64 + lane];
        }
    }
    __syncthreads();
  }
}

// Does this wave hold final accumulators (all waves, or wave 0 of a K-split tile)?
template <typename Cfg>
__device__ __forceinline__ bool holds_result() {
  return Cfg::WGK == 1 || (threadIdx.x >> 6) == 0;
}

// Epilogue: C = alpha * acc + beta * C, scattered through the C tables.
template <typename S2, typename R, typename Cfg>
__device__ __forceinline__ void gemm_store(const TileView& p, int64_t tm, int64_t tn, int64_t z,
                                           const TileAcc<R, Cfg>& acc) {
  if (!holds_result<Cfg>()) return;
  S2* __restrict__ C = static_cast<S2*>(p.C);
  const int lane = threadIdx.x & 63;
  const int wmn = (threadIdx.x >> 6) / Cfg::WGK;
  const int wm = wmn % Cfg::WGM, wn = wmn / Cfg::WGM;
  const int64_t m0 = tm * Cfg::BM, n0 = tn * Cfg::BN;
  const int64_t offCz = p.t.bC[z];
  const double ar = p.alpha_re, ai = p.alpha_im, br = p.beta_re, bi = p.beta_im;
  const bool use_beta = (br != 0.0 || bi != 0.0);
#pragma unroll
  for (int j = 0; j < Cfg::TN; ++j) {
    const int64_t col = n0 + (wn * Cfg::TN + j) * 16 + (lane & 15);
    if (col >= p.N) continue;
    const int64_t ccol = p.t.colC[col] + offCz;
#pragma unroll
    for (int i = 0; i < Cfg::TM; ++i) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        // D layout: FP32 16x16x4 keeps rows 4*(l/16)+r in lane l; FP64
        // 16x16x4 interleaves them as rows 4*r + l/16.
        const int rr = (sizeof(R) == 8) ? (r * 4 + (lane >> 4)) : ((lane >> 4) * 4 + r);
        const int64_t row = m0 + (wm * Cfg::TM + i) * 16 + rr;
        if (row >= p.M) continue;
        const int64_t off = p.t.rowC[row] + ccol;
        const double cr = (double)acc.r[i][j][r], ci = (double)acc.i[i][j][r];
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

// Split-K partials: a result-holding thread stores its registers at
// ((wave * TM + i) * TN + j) * 4 + r) * 64 + lane, as (re, im) pairs; the
// same thread of the reducing workgroup reads them back from the same places.
template <typename R, typename Cfg>
__device__ __forceinline__ void partial_store(R* __restrict__ part, const TileAcc<R, Cfg>& acc) {
  if (!holds_result<Cfg>()) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < Cfg::TM; ++i)
#pragma unroll
    for (int j = 0; j < Cfg::TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t e = ((((int64_t)wave * Cfg::TM + i) * Cfg::TN + j) * 4 + r) * 64 + lane;
        part[2 * e] = acc.r[i][j][r];
        part[2 * e + 1] = acc.i[i][j][r];
      }
}

template <typename R, typename Cfg>
__device__ __forceinline__ void partial_sum(const R* __restrict__ part, int nchunks, TileAcc<R, Cfg>& acc) {
  if (!holds_result<Cfg>()) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int64_t TILE = (int64_t)Cfg::BM * Cfg::BN;
#pragma unroll
  for (int i = 0; i < Cfg::TM; ++i)
#pragma unroll
    for (int j = 0; j < Cfg::TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t e = ((((int64_t)wave * Cfg::TM + i) * Cfg::TN + j) * 4 + r) * 64 + lane;
        R sr = 0, si = 0;
        for (int c = 0; c < nchunks; ++c) {  // fixed order: deterministic
          sr += part[2 * (c * TILE + e)];
          si += part[2 * (c * TILE + e) + 1];
        }
        acc.r[i][j][r] = sr;
        acc.i[i][j][r] = si;
      }
}

// One Cfg::BM x Cfg::BN output tile (tm, tn) of batch entry z over the whole K range.
template <typename S2, typename R, typename Cfg>
__device__ __forceinline__ void gemm_tile(const TileView& p, int64_t tm, int64_t tn, int64_t z, GemmSmem<R>& sm) {
  TileAcc<R, Cfg> acc;
  gemm_accumulate<S2, R, Cfg>(p, tm, tn, z, 0, p.K, sm, acc);
  gemm_store<S2, R, Cfg>(p, tm, tn, z, acc);
}

// Map a physical workgroup id to a logical tile so that the tiles one XCD
// receives (ids congruent mod 8 are dispatched to the same XCD) are
// neighbours in the tile grid and share A/B panels in that XCD's L2.
__device__ __forceinline__ int64_t xcd_remap(int64_t pid, int64_t total) {
  constexpr int64_t NXCD = 8;
  if (total % NXCD != 0) return pid;
  return (pid % NXCD) * (total / NXCD) + pid / NXCD;
}

template <typename S2, typename R, typename Cfg>
__global__ __launch_bounds__(256) void gemm_kernel(TileView p, int64_t tiles_m, int64_t tiles_n, int64_t batch) {
  __shared__ GemmSmem<R> sm;
  const int64_t per_batch = tiles_m * tiles_n;
  const int64_t total = per_batch * batch;
  const int64_t lid = xcd_remap((int64_t)blockIdx.x, total);
  const int64_t z = lid / per_batch;
  const int64_t t = lid - z * per_batch;
  gemm_tile<S2, R, Cfg>(p, t % tiles_m, t / tiles_m, z, sm);
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
  // 64 x 64 tiles, or 32 x 32 when 64 x 64 would leave most of the 256 CUs idle
  // (a 512 x 512 product: 64 workgroups against 256)
  const bool small = ((g.M + 63) / 64) * ((g.N + 63) / 64) * g.batch < 128;
  const int64_t bt = small ? 32 : 64;
  const int64_t tm = (g.M + bt - 1) / bt, tn = (g.N + bt - 1) / bt;
  const int64_t total = tm * tn * g.batch;
  if (total <= 0) return TENET_OK;
  if (total > 0x7fffffffLL) {
    set_error("contraction too large for one launch");
    return TENET_ERR_UNSUPPORTED;
  }
  if (small)
    hipLaunchKernelGGL((gemm_kernel<S2, R, Tile32>), dim3((unsigned)total), dim3(256), 0, s, p, tm, tn, g.batch);
  else
    hipLaunchKernelGGL((gemm_kernel<S2, R, Tile64>), dim3((unsigned)total), dim3(256), 0, s, p, tm, tn, g.batch);
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
    case TENET_PREC_SPLIT_I8:
    case TENET_PREC_SPLIT_I8_FAST:
      return launch_split8_gemm(g, TENET_C128, prec, s, ws, ws_bytes);
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
    case TENET_PREC_SPLIT_I8:
    case TENET_PREC_SPLIT_I8_FAST:
      return launch_split8_gemm(g, TENET_C64, prec, s, ws, ws_bytes);
  }
  set_error("unknown precision");
  return TENET_ERR_ARG;
}

// ---- grouped persistent launch ------------------------------------------------
// Work units (problem, tile, K chunk) are sorted by decreasing cost on the
// host; persistent workgroups take the next unit from a device counter, so
// the largest units start first and the tail is short (LPT order).  A tile
// whose K range is cut into chunks stores each chunk's partial tile; the
// workgroup that finishes a tile's last chunk (per-tile arrival counter) sums
// the partials in chunk order and writes C, so results do not depend on the
// schedule.
template <typename S2, typename R, typename Cfg>
__device__ __forceinline__ void grouped_unit(const GroupedProblem& q, const uint64_t* __restrict__ ptrs,
                                             const GroupedUnit& u, R* __restrict__ partials,
                                             uint32_t* __restrict__ tile_ctr, GemmSmem<R>& sm, int* bcast) {
  TileView p;
  p.A = (const void*)ptrs[3 * u.prob];
  p.B = (const void*)ptrs[3 * u.prob + 1];
  p.C = (void*)ptrs[3 * u.prob + 2];
  p.t = q.t;
  p.M = q.M;
  p.N = q.N;
  p.K = q.K;
  p.a_mfast = q.a_mfast;
  p.b_nfast = q.b_nfast;
  p.conjA = q.conjA;
  p.conjB = q.conjB;
  p.alpha_re = 1.0;
  p.alpha_im = 0.0;
  p.beta_re = 0.0;
  p.beta_im = 0.0;
  const int64_t per_batch = (int64_t)q.tiles_m * q.tiles_n;
  const int64_t z = u.tile / per_batch;
  const int64_t t = u.tile - z * per_batch;
  const int64_t tm = t % q.tiles_m, tn = t / q.tiles_m;
  const int64_t k_lo = (int64_t)u.chunk * q.kc;
  const int64_t k_hi = std::min<int64_t>(q.K, k_lo + q.kc);
  TileAcc<R, Cfg> acc;
  if (q.conjA || q.conjB)
    gemm_accumulate<S2, R, Cfg, true>(p, tm, tn, z, k_lo, k_hi, sm, acc);
  else
    gemm_accumulate<S2, R, Cfg, false>(p, tm, tn, z, k_lo, k_hi, sm, acc);
  if (q.kchunks == 1) {
    gemm_store<S2, R, Cfg>(p, tm, tn, z, acc);
  } else {
    constexpr int64_t TILE = (int64_t)Cfg::BM * Cfg::BN;
    R* part = partials + 2 * (q.part_base + u.tile * q.kchunks * TILE);
    partial_store<R, Cfg>(part + 2 * (int64_t)u.chunk * TILE, acc);
    __threadfence();  // release this thread's partial stores at device scope
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint32_t prev = __hip_atomic_fetch_add(&tile_ctr[q.ctr_base + u.tile], 1u, __ATOMIC_ACQ_REL,
                                                   __HIP_MEMORY_SCOPE_AGENT);
      *bcast = (prev + 1 == (uint32_t)q.kchunks) ? 1 : 0;
    }
    __syncthreads();
    // wave-uniform by construction (one LDS word after a barrier); readfirstlane
    // lets the compiler branch on it with scalar control flow
    const int last = __builtin_amdgcn_readfirstlane(*bcast);
    __syncthreads();  // bcast is reused by the next unit
    if (last) {
      __threadfence();  // acquire the other chunks' partials
      partial_sum<R, Cfg>(part, q.kchunks, acc);
      gemm_store<S2, R, Cfg>(p, tm, tn, z, acc);
      if (threadIdx.x == 0) tile_ctr[q.ctr_base + u.tile] = 0u;  // ready for the next launch
    }
  }
}

// A GROUPED_SMALLK unit: output elements [tile * SMALLK_ELEMS, +SMALLK_ELEMS)
// of the flattened (m, n, z) space, the class q.order varying fastest (the
// one with C's smallest stride, so consecutive threads write consecutive
// addresses); each thread sums its K <= 8 products directly.  The element
// index is decoded in 32-bit arithmetic when the problem's element count
// fits (a 64-bit division is a long software sequence on the GPU; the
// Vidal layer's gate contraction spent most of its time there).
template <typename S2, typename R, typename IT>
__device__ __forceinline__ void smallk_elems(const GroupedProblem& q, const S2* __restrict__ A,
                                             const S2* __restrict__ B, S2* __restrict__ C, IT e0, IT e1, IT ext0,
                                             IT ext1) {
  const R sa = q.conjA ? (R)-1 : (R)1, sb = q.conjB ? (R)-1 : (R)1;
  for (IT e = e0 + (IT)threadIdx.x; e < e1; e += (IT)blockDim.x) {
    const IT i0 = e % ext0, rest = e / ext0;
    const IT i1 = rest % ext1, i2 = rest / ext1;
    int64_t m, n, z;
    if (q.order == 0) {
      m = i0;
      n = i1;
      z = i2;
    } else if (q.order == 1) {
      n = i0;
      m = i1;
      z = i2;
    } else {
      z = i0;
      m = i1;
      n = i2;
    }
    const int64_t oa = q.t.rowA[m] + q.t.bA[z], ob = q.t.colB[n] + q.t.bB[z];
    R cr = 0, ci = 0;
    // K <= 8, unrolled: the element loop stays the only loop of the unit
    // (the persistent loop around it must keep a single nested level)
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (k < q.K) {
        const S2 a = A[oa + q.t.kA[k]];
        const S2 b = B[ob + q.t.kB[k]];
        const R ar = (R)a.x, ai = sa * (R)a.y, br = (R)b.x, bi = sb * (R)b.y;
        cr += ar * br - ai * bi;
        ci += ar * bi + ai * br;
      }
    }
    S2 o;
    o.x = cr;
    o.y = ci;
    C[q.t.rowC[m] + q.t.colC[n] + q.t.bC[z]] = o;
  }
}

template <typename S2, typename R>
__device__ __forceinline__ void grouped_smallk_unit(const GroupedProblem& q, const uint64_t* __restrict__ ptrs,
                                                    const GroupedUnit& u) {
  const S2* __restrict__ A = reinterpret_cast<const S2*>(ptrs[3 * u.prob]);
  const S2* __restrict__ B = reinterpret_cast<const S2*>(ptrs[3 * u.prob + 1]);
  S2* __restrict__ C = reinterpret_cast<S2*>(ptrs[3 * u.prob + 2]);
  const int64_t total = q.M * q.N * q.batch;
  const int64_t e0 = u.tile * SMALLK_ELEMS;
  const int64_t e1 = std::min<int64_t>(total, e0 + SMALLK_ELEMS);
  // extents in decode order: fastest, middle, slowest
  const int64_t ext0 = q.order == 0 ? q.M : (q.order == 1 ? q.N : q.batch);
  const int64_t ext1 = q.order == 0 ? q.N : q.M;
  if (total <= (int64_t)INT32_MAX)
    smallk_elems<S2, R, uint32_t>(q, A, B, C, (uint32_t)e0, (uint32_t)e1, (uint32_t)ext0, (uint32_t)ext1);
  else
    smallk_elems<S2, R, int64_t>(q, A, B, C, e0, e1, ext0, ext1);
}

template <typename S2, typename R>
__device__ __forceinline__ void grouped_gemm_body(const GroupedProblem* __restrict__ probs,
                                                  const GroupedUnit* __restrict__ units, int64_t nunits,
                                                  const uint64_t* __restrict__ ptrs, R* __restrict__ partials,
                                                  uint32_t* __restrict__ counters) {
  __shared__ GemmSmem<R> sm;
  __shared__ uint32_t next;
  __shared__ int bcast;
  uint32_t* queue = counters;          // counters[0]: next unit to take
  uint32_t* tile_ctr = counters + 64;  // per split tile arrival counts
  // Every branch below depends only on values that are uniform across the
  // workgroup and are made scalar with readfirstlane, so the loop compiles to
  // scalar control flow: each iteration takes exactly one unit from the queue
  // and every workgroup leaves once the queue is drained.
  for (;;) {
    // The thread index is re-read through an opaque copy every iteration.
    // Otherwise the optimiser sees "thread 0" as loop-invariant and threads
    // the loop latch on it (thread 0 back to the queue fetch, every other
    // thread straight back to the first barrier), which splits the barrier
    // instances and never fetches a new unit for the other lanes.
    unsigned tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    if (tid == 0) next = __hip_atomic_fetch_add(queue, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int64_t ui = (int64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)next);
    __syncthreads();
    if (ui >= nunits) break;
    const GroupedUnit u = units[ui];
    const GroupedProblem& q = probs[u.prob];
    switch (q.shape) {
      case GROUPED_SMALLK:
        grouped_smallk_unit<S2, R>(q, ptrs, u);
        break;
      case GROUPED_TILE_16:
        grouped_unit<S2, R, Tile16>(q, ptrs, u, partials, tile_ctr, sm, &bcast);
        break;
      case GROUPED_TILE_32:
        grouped_unit<S2, R, Tile32>(q, ptrs, u, partials, tile_ctr, sm, &bcast);
        break;
      default:
        grouped_unit<S2, R, Tile64>(q, ptrs, u, partials, tile_ctr, sm, &bcast);
        break;
    }
  }
}

// complex64: two resident workgroups per CU fit without spills; complex128
// keeps one (its FP64 accumulators double the register need)
__global__ __launch_bounds__(256, 2) void grouped_gemm_c64_kernel(const GroupedProblem* __restrict__ probs,
                                                                  const GroupedUnit* __restrict__ units,
                                                                  int64_t nunits, const uint64_t* __restrict__ ptrs,
                                                                  float* __restrict__ partials,
                                                                  uint32_t* __restrict__ counters) {
  grouped_gemm_body<float2, float>(probs, units, nunits, ptrs, partials, counters);
}
__global__ __launch_bounds__(256) void grouped_gemm_c128_kernel(const GroupedProblem* __restrict__ probs,
                                                                const GroupedUnit* __restrict__ units, int64_t nunits,
                                                                const uint64_t* __restrict__ ptrs,
                                                                double* __restrict__ partials,
                                                                uint32_t* __restrict__ counters) {
  grouped_gemm_body<double2, double>(probs, units, nunits, ptrs, partials, counters);
}

int grouped_occupancy(int dtype) {
  int blocks = 0;
  hipError_t e;
  if (dtype == TENET_C128)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, grouped_gemm_c128_kernel, 256, 0);
  else
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, grouped_gemm_c64_kernel, 256, 0);
  if (e != hipSuccess || blocks < 1) return 1;
  return std::min(blocks, 8);
}

int launch_grouped_gemm(const GroupedLaunch& g, hipStream_t s) {
  if (g.nunits <= 0) return TENET_OK;
  // the queue head is cleared by a kernel (graph-safe); tile counters are
  // left at zero by the reducing workgroups and cleared at plan creation
  {
    const int st = launch_fill_zero(g.counters, sizeof(uint32_t), s);
    if (st != TENET_OK) return st;
  }
  const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(g.nunits, g.grid));
  if (g.dtype == TENET_C128)
    hipLaunchKernelGGL(grouped_gemm_c128_kernel, dim3((unsigned)grid), dim3(256), 0, s, g.probs,
                       g.units, g.nunits, g.ptrs, (double*)g.partials, g.counters);
  else
    hipLaunchKernelGGL(grouped_gemm_c64_kernel, dim3((unsigned)grid), dim3(256), 0, s, g.probs,
                       g.units, g.nunits, g.ptrs, (float*)g.partials, g.counters);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

}  // namespace tenet
