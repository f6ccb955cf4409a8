// Split-precision contraction on int8 MFMA (v_mfma_i32_16x16x64_i8, gfx950).
//
// The same error-free scheme as k_split.hip (Ozaki, Ogita, Oishi, Rump,
// Numer. Algorithms 59 (2012): operands cut into integer slices whose
// products are exact), on the 8-bit integer matrix cores with exact int32
// accumulation instead of bf16 / fp16 with FP32 accumulation:
//
//  * every complex row of A (over k) and column of B gets a power-of-two
//    scale 2^e with max(|Re|, |Im|) < 2^e; the scaled value x = v 2^(7-e)
//    lies in (-128, 128) and is cut into S signed 7-bit slices by
//    truncation toward zero, s_1 = trunc(x), x_1 = (x - s_1) 128, ...,
//    so |s_t| <= 127 and the remainder after S slices is below 2^-7S of the
//    scale;
//  * complex arithmetic as one real GEMM: A' has two rows per complex row m,
//    [Re A_m, -Im A_m] and [Im A_m, Re A_m] over k' in [0, 2K), and B' the
//    column [Re B_n; Im B_n], so A' B' holds Re C_m,n and Im C_m,n in rows
//    2m and 2m+1;
//  * products of slices are integers below 2^14; a level (slice pairs with
//    s + t = l, at most S of them) sums exactly in int32 over blocks of
//    8192 k' (S * 8192 * 127^2 < 2^31), blocks and levels are combined in
//    FP64:  C = 2^(e_m + f_n - 14) sum_l 2^(-7 l) sum_{s+t=l} A'_s B'_t,
//    keeping the S(S+1)/2 pairs with s + t <= S - 1.
//
// Accuracy: slice remainders and dropped levels are ~2^-7S relative to the
// row / column maxima: 2^-49 for S = 7 (FP64 grade), 2^-35 for S = 5, 2^-28
// for S = 4 (complex64 results).
//
// GEMM tile: 128 stacked rows (64 complex rows; 64 stacked rows for seven
// slices) x 64 columns per 256-thread workgroup, K step 64 (one MFMA), every
// slice of both operand tiles staged in LDS once per step and reused by all
// S(S+1)/2 slice pairs; wave w owns its two (one) 16-row blocks and all four
// 16-column blocks, with one int32 accumulator set per level.  The next step's tiles are
// loaded into registers while the current one is multiplied.
#include "../../include/tenet_hip.h"
#include "internal.h"

namespace tenet {

typedef int i4v __attribute__((ext_vector_type(4)));

constexpr int I8_BM = 128;   // stacked-row padding (a multiple of every tile height)
constexpr int I8_BN = 64;    // columns per tile
constexpr int I8_KS = 64;    // k' per step (one MFMA)
constexpr int I8_KB = 8192;  // k' per exact int32 block
constexpr int I8_LDS_ROW = I8_KS + 16;  // padded row (bytes): spreads 16-byte reads over the banks

struct I8Layout {
  int64_t Mp2, Np, Kp2;  // padded stacked rows, columns, 2K
  size_t a_bytes, b_bytes, ea_bytes, eb_bytes;
};

static I8Layout i8_layout(const GemmLaunch& g, int S) {
  I8Layout L;
  L.Mp2 = (2 * g.M + I8_BM - 1) / I8_BM * I8_BM;
  L.Np = (g.N + I8_BN - 1) / I8_BN * I8_BN;
  L.Kp2 = (2 * g.K + I8_KS - 1) / I8_KS * I8_KS;
  L.a_bytes = (size_t)g.batch * S * L.Mp2 * L.Kp2;
  L.b_bytes = (size_t)g.batch * S * L.Np * L.Kp2;
  L.ea_bytes = (size_t)g.batch * (L.Mp2 / 2) * 4;
  L.eb_bytes = (size_t)g.batch * L.Np * 4;
  return L;
}

static size_t al256(size_t x) { return (x + 255) / 256 * 256; }

int split8_slices(int dtype, int prec) {
  if (dtype == TENET_C64) return 4;
  return prec == TENET_PREC_SPLIT_I8_FAST ? 5 : 7;
}

size_t split8_workspace_bytes(const GemmLaunch& g, int dtype, int prec) {
  const I8Layout L = i8_layout(g, split8_slices(dtype, prec));
  return al256(L.a_bytes) + al256(L.b_bytes) + al256(L.ea_bytes) + al256(L.eb_bytes) + 1024;
}

// One wave per complex row (A: row m over k; B: column n over k), lanes over
// k.  side 0 (A): stacked rows 2m = [Re, -Im], 2m+1 = [Im, Re] over k';
// side 1 (B): the column [Re; Im].  Layout [z][slice][row'][Kp2] int8,
// exps[z][row] = e.
template <typename S2, int S>
__global__ __launch_bounds__(256) void split8_pack_kernel(const S2* __restrict__ src, const int64_t* __restrict__ tfix,
                                                          const int64_t* __restrict__ tk,
                                                          const int64_t* __restrict__ tb, int64_t R, int64_t K,
                                                          int64_t Rp, int64_t Kp2, int64_t batch, int conj, int side,
                                                          int8_t* __restrict__ out, int* __restrict__ exps) {
  const int lane = threadIdx.x & 63;
  const int64_t row_global = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row_global >= Rp * batch) return;
  const int64_t z = row_global / Rp, row = row_global % Rp;
  const bool live = row < R;
  const int64_t base = live ? tb[z] + tfix[row] : 0;
  double mx = 0.0;
  if (live)
    for (int64_t k = lane; k < K; k += 64) {
      const S2 v = src[base + tk[k]];
      mx = fmax(mx, fmax(fabs((double)v.x), fabs((double)v.y)));
    }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
  int e = 0;
  if (mx > 0.0) (void)frexp(mx, &e);  // mx = f 2^e, f in [0.5, 1): |v| 2^-e < 1
  if (lane == 0) exps[z * Rp + row] = e;
  const int64_t rows_out = side == 0 ? 2 * Rp : Rp;
  const size_t plane = (size_t)rows_out * Kp2;
  int8_t* base_out = out + (size_t)z * S * plane;
  for (int64_t k = lane; k < Kp2 / 2; k += 64) {
    double xr = 0.0, xi = 0.0;
    if (live && k < K) {
      const S2 v = src[base + tk[k]];
      xr = ldexp((double)v.x, 7 - e);
      xi = ldexp(conj ? -(double)v.y : (double)v.y, 7 - e);
    }
    const int64_t kh = Kp2 / 2;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const double sr = trunc(xr), si = trunc(xi);
      xr = (xr - sr) * 128.0;
      xi = (xi - si) * 128.0;
      const int8_t ar = (int8_t)(int)sr, ai = (int8_t)(int)si, an = (int8_t)(int)(-si);
      int8_t* o = base_out + (size_t)s * plane;
      if (side == 0) {
        o[(size_t)(2 * row) * Kp2 + k] = ar;
        o[(size_t)(2 * row) * Kp2 + kh + k] = an;
        o[(size_t)(2 * row + 1) * Kp2 + k] = ai;
        o[(size_t)(2 * row + 1) * Kp2 + kh + k] = ar;
      } else {
        o[(size_t)row * Kp2 + k] = ar;
        o[(size_t)row * Kp2 + kh + k] = ai;
      }
    }
  }
}

struct Split8Args {
  const int8_t* Ap;
  const int8_t* Bp;
  const int* ea;
  const int* eb;
  void* C;
  const int64_t* rowC;
  const int64_t* colC;
  const int64_t* bC;
  int64_t M, N, Mp2, Np, Kp2, batch;
  double alpha_re, alpha_im, beta_re, beta_im;
  int64_t tiles_m, tiles_n;
};

template <int S, int RB>
struct I8Smem {
  int8_t a[S][64 * RB][I8_LDS_ROW];
  int8_t b[S][I8_BN][I8_LDS_ROW];
};

// RB: 16-row blocks per wave (the tile holds 64 RB stacked rows); 2 up to
// five slices, 1 for seven (the accumulators of 2 x 4 blocks x 7 levels
// would not fit the register file)
template <typename S2, int S, int RB>
__global__ __launch_bounds__(256) void split8_gemm_kernel(Split8Args a) {
  __shared__ I8Smem<S, RB> sm;
  constexpr int BM = 64 * RB;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t per = a.tiles_m * a.tiles_n;
  const int64_t tile = blockIdx.x;
  const int64_t z = tile / per, t = tile % per;
  const int64_t m0 = (t % a.tiles_m) * BM, n0 = (t / a.tiles_m) * I8_BN;
  const size_t aplane = (size_t)a.Mp2 * a.Kp2, bplane = (size_t)a.Np * a.Kp2;
  const int8_t* Az = a.Ap + (size_t)z * S * aplane;
  const int8_t* Bz = a.Bp + (size_t)z * S * bplane;
  // staging: per slice the A tile is 64 RB rows x 4 chunks of 16 bytes (RB
  // per thread) and the B tile 64 rows x 4 chunks (1 per thread)
  const int sra = tid >> 2, sc = tid & 3;  // A rows sra + 64 q, B row sra; chunk sc
  i4v ra[S][RB], rb[S];
  auto gload = [&](int64_t k0) {
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int8_t* pa = Az + (size_t)s * aplane + (size_t)(m0 + sra) * a.Kp2 + k0 + 16 * sc;
#pragma unroll
      for (int q = 0; q < RB; ++q) ra[s][q] = *reinterpret_cast<const i4v*>(pa + (size_t)(64 * q) * a.Kp2);
      const int8_t* pb = Bz + (size_t)s * bplane + (size_t)(n0 + sra) * a.Kp2 + k0 + 16 * sc;
      rb[s] = *reinterpret_cast<const i4v*>(pb);
    }
  };
  auto sstore = [&]() {
#pragma unroll
    for (int s = 0; s < S; ++s) {
#pragma unroll
      for (int q = 0; q < RB; ++q) *reinterpret_cast<i4v*>(&sm.a[s][sra + 64 * q][16 * sc]) = ra[s][q];
      *reinterpret_cast<i4v*>(&sm.b[s][sra][16 * sc]) = rb[s];
    }
  };
  // accumulators: level l, row block i (RB), column block j (4)
  i4v acc[S][RB][4];
  double tot[RB][4][4];
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) tot[i][j][r] = 0.0;
  const int fr = lane & 15, fk = 16 * (lane >> 4);  // operand fragment: row / column, k offset
  gload(0);
  for (int64_t kb = 0; kb < a.Kp2; kb += I8_KB) {
    const int64_t ke = kb + I8_KB < a.Kp2 ? kb + I8_KB : a.Kp2;
#pragma unroll
    for (int l = 0; l < S; ++l)
#pragma unroll
      for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[l][i][j] = i4v{0, 0, 0, 0};
    for (int64_t k0 = kb; k0 < ke; k0 += I8_KS) {
      __syncthreads();  // the previous step's readers are done with the tiles
      sstore();
      __syncthreads();
      if (k0 + I8_KS < a.Kp2) gload(k0 + I8_KS);
      // every A fragment of the step once, each B slice's fragments once
      i4v af[S][RB];
#pragma unroll
      for (int sa = 0; sa < S; ++sa)
#pragma unroll
        for (int i = 0; i < RB; ++i)
          af[sa][i] = *reinterpret_cast<const i4v*>(&sm.a[sa][16 * RB * wave + 16 * i + fr][fk]);
#pragma unroll
      for (int sb = 0; sb < S; ++sb) {
        i4v bf[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) bf[j] = *reinterpret_cast<const i4v*>(&sm.b[sb][16 * j + fr][fk]);
#pragma unroll
        for (int sa = 0; sa + sb < S; ++sa)
#pragma unroll
          for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int i = 0; i < RB; ++i)
              acc[sa + sb][i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[sa][i], bf[j], acc[sa + sb][i][j], 0, 0, 0);
      }
    }
    // fold the exact block into FP64: sum_l 2^-7l level_l
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          double v = 0.0, w = 1.0;
#pragma unroll
          for (int l = 0; l < S; ++l) {
            v += w * (double)acc[l][i][j][r];
            w *= 0x1p-7;
          }
          tot[i][j][r] += v;
        }
  }
  // epilogue: accumulator r of lane l holds stacked row 4 (l / 16) + r of the
  // block, column l % 16; rows (4q, 4q + 1) are (Re, Im) of complex row 2q,
  // rows (4q + 2, 4q + 3) of complex row 2q + 1
  S2* C = static_cast<S2*>(a.C);
  const int64_t offCz = a.bC[z];
  const bool use_beta = (a.beta_re != 0.0 || a.beta_im != 0.0);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int64_t col = n0 + 16 * j + (lane & 15);
    if (col >= a.N) continue;
    const int fcol = a.eb[z * a.Np + col];
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int64_t srow = m0 + 16 * RB * wave + 16 * i + 4 * (lane >> 4) + 2 * h;  // stacked row of Re
        const int64_t row = srow / 2;
        if (row >= a.M) continue;
        const int sc2 = a.ea[z * (a.Mp2 / 2) + row] + fcol - 14;
        const double cr = ldexp(tot[i][j][2 * h], sc2), ci = ldexp(tot[i][j][2 * h + 1], sc2);
        double outr = a.alpha_re * cr - a.alpha_im * ci;
        double outi = a.alpha_re * ci + a.alpha_im * cr;
        const int64_t off = a.rowC[row] + a.colC[col] + offCz;
        if (use_beta) {
          const S2 old = C[off];
          outr += a.beta_re * (double)old.x - a.beta_im * (double)old.y;
          outi += a.beta_re * (double)old.y + a.beta_im * (double)old.x;
        }
        S2 o;
        o.x = outr;
        o.y = outi;
        C[off] = o;
      }
  }
}

template <typename S2, int S>
static int run_split8(const GemmLaunch& g, hipStream_t s, void* ws) {
  constexpr int RB = S <= 5 ? 2 : 1;
  const I8Layout L = i8_layout(g, S);
  char* w = static_cast<char*>(ws);
  int8_t* Ap = reinterpret_cast<int8_t*>(w);
  w += al256(L.a_bytes);
  int8_t* Bp = reinterpret_cast<int8_t*>(w);
  w += al256(L.b_bytes);
  int* ea = reinterpret_cast<int*>(w);
  w += al256(L.ea_bytes);
  int* eb = reinterpret_cast<int*>(w);
  {
    const int64_t rows = (L.Mp2 / 2) * g.batch;
    hipLaunchKernelGGL((split8_pack_kernel<S2, S>), dim3((unsigned)((rows + 3) / 4)), dim3(256), 0, s,
                       static_cast<const S2*>(g.A), g.t.rowA, g.t.kA, g.t.bA, g.M, g.K, L.Mp2 / 2, L.Kp2, g.batch,
                       g.conjA, 0, Ap, ea);
    TENET_HIP_CHECK(hipGetLastError());
  }
  {
    const int64_t rows = L.Np * g.batch;
    hipLaunchKernelGGL((split8_pack_kernel<S2, S>), dim3((unsigned)((rows + 3) / 4)), dim3(256), 0, s,
                       static_cast<const S2*>(g.B), g.t.colB, g.t.kB, g.t.bB, g.N, g.K, L.Np, L.Kp2, g.batch, g.conjB,
                       1, Bp, eb);
    TENET_HIP_CHECK(hipGetLastError());
  }
  Split8Args a;
  a.Ap = Ap;
  a.Bp = Bp;
  a.ea = ea;
  a.eb = eb;
  a.C = g.C;
  a.rowC = g.t.rowC;
  a.colC = g.t.colC;
  a.bC = g.t.bC;
  a.M = g.M;
  a.N = g.N;
  a.Mp2 = L.Mp2;
  a.Np = L.Np;
  a.Kp2 = L.Kp2;
  a.batch = g.batch;
  a.alpha_re = g.alpha_re;
  a.alpha_im = g.alpha_im;
  a.beta_re = g.beta_re;
  a.beta_im = g.beta_im;
  a.tiles_m = L.Mp2 / (64 * RB);
  a.tiles_n = L.Np / I8_BN;
  const int64_t total = a.tiles_m * a.tiles_n * g.batch;
  if (total > 0x7fffffffLL) {
    set_error("split contraction too large for one launch");
    return TENET_ERR_UNSUPPORTED;
  }
  hipLaunchKernelGGL((split8_gemm_kernel<S2, S, RB>), dim3((unsigned)total), dim3(256), 0, s, a);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

int launch_split8_gemm(const GemmLaunch& g, int dtype, int prec, hipStream_t s, void* ws, size_t ws_bytes) {
  if (!ws || ws_bytes < split8_workspace_bytes(g, dtype, prec)) {
    set_error("int8 split contraction: workspace too small");
    return TENET_ERR_OOM;
  }
  if (g.K == 0) {
    set_error("split contraction with an empty contracted index");
    return TENET_ERR_UNSUPPORTED;
  }
  if (dtype == TENET_C128)
    return prec == TENET_PREC_SPLIT_I8_FAST ? run_split8<double2, 5>(g, s, ws) : run_split8<double2, 7>(g, s, ws);
  return run_split8<float2, 4>(g, s, ws);
}

}  // namespace tenet
