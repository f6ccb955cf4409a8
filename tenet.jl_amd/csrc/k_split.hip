// Split-precision contraction on bf16 / fp16 MFMA with error-free FP32
// accumulation (an Ozaki-style scheme: Ozaki, Ogita, Oishi, Rump, Numer.
// Algorithms 59 (2012); applied to low-precision matrix units as in Ootomo &
// Yokota, IJHPCA 2022, but with exact integer slices instead of a residual).
//
// Every row of A (over k) and every column of B gets a power-of-two scale
// 2^e with max(|Re|, |Im|) < 2^e.  The scaled value x = v 2^-e in (-1, 1) is
// cut into S 8-bit integer slices (S = 6 for complex128 results, 3 for
// complex64),
//     x 2^8 = s1 + 2^-8 s2 + ... + 2^-8(S-1) sS + O(2^-(8S+1)),  |s_i| <= 256,
// each exact in bf16 (8 significant bits) and fp16.  A product of two slices
// is an integer below 2^16 in magnitude, so a K block of 128 complex terms
// (256 real products per accumulator) sums exactly in FP32 (< 2^24): the
// two leading levels (i + j = 2, 3) are error free, later levels carry
// weights <= 2^-16.  Each level is accumulated separately and all K blocks
// are summed in FP64:
//     C = 2^(e_m + f_n - 16) sum_{l} 2^(-8 (l-2)) sum_{i+j=l} S_i T_j,
// keeping the S(S+1)/2 products with i + j <= S + 1.  The dropped terms and
// slice remainders are ~2^-(8S+1) relative to the row / column maxima:
// 2^-49 (FP64 grade, needed where a TEBD sweep amplifies contraction error
// by ~1e5) for S = 6, 2^-25 (FP32 grade) for S = 3.
#include "../../include/tenet_hip.h"
#include "internal.h"

namespace tenet {

typedef float f4s __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef unsigned int u4 __attribute__((ext_vector_type(4)));

constexpr int NSL_MAX = 6;  // slices for complex128 results
constexpr int SKB = 128;  // exact-accumulation K block
constexpr int SKP = 32;   // MFMA K

struct SplitLayout {
  int64_t Mp, Np, Kp;  // padded sizes
  size_t a_bytes, b_bytes, ea_bytes, eb_bytes;
};

static SplitLayout split_layout(const GemmLaunch& g) {
  SplitLayout L;
  L.Mp = (g.M + 63) / 64 * 64;
  L.Np = (g.N + 31) / 32 * 32;
  L.Kp = (g.K + SKB - 1) / SKB * SKB;
  L.a_bytes = (size_t)g.batch * NSL_MAX * 2 * L.Mp * L.Kp * 2;
  L.b_bytes = (size_t)g.batch * NSL_MAX * 2 * L.Np * L.Kp * 2;
  L.ea_bytes = (size_t)g.batch * L.Mp * 4;
  L.eb_bytes = (size_t)g.batch * L.Np * 4;
  return L;
}

static size_t align256(size_t x) { return (x + 255) / 256 * 256; }

size_t gemm_workspace_bytes(const GemmLaunch& g, int dtype, int prec) {
  if (prec == TENET_PREC_SPLIT_I8 || prec == TENET_PREC_SPLIT_I8_FAST) return split8_workspace_bytes(g, dtype, prec);
  if (prec != TENET_PREC_SPLIT_BF16 && prec != TENET_PREC_SPLIT_FP16) return 0;
  const SplitLayout L = split_layout(g);
  return align256(L.a_bytes) + align256(L.b_bytes) + align256(L.ea_bytes) + align256(L.eb_bytes) + 1024;
}

template <typename T>
__device__ __forceinline__ unsigned short to_bits(double v);
template <>
__device__ __forceinline__ unsigned short to_bits<__bf16>(double v) {
  const __bf16 h = (__bf16)(float)v;  // exact: |v| <= 256 is an integer
  return __builtin_bit_cast(unsigned short, h);
}
template <>
__device__ __forceinline__ unsigned short to_bits<_Float16>(double v) {
  const _Float16 h = (_Float16)(float)v;
  return __builtin_bit_cast(unsigned short, h);
}

// One wave per packed row (A: row m over k; B: column n over k).  out layout:
// [z][slice][part][row][Kp] of 16-bit codes; exps[z][row] = e.
template <typename S2, typename T, int NSL>
__global__ __launch_bounds__(256) void split_pack_kernel(const S2* __restrict__ src, const int64_t* __restrict__ tfix,
                                                         const int64_t* __restrict__ tk,
                                                         const int64_t* __restrict__ tb, int64_t R, int64_t K,
                                                         int64_t Rp, int64_t Kp, int64_t batch, int conj,
                                                         unsigned short* __restrict__ out, int* __restrict__ exps) {
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
  const size_t plane = (size_t)Rp * Kp;
  for (int64_t k = lane; k < Kp; k += 64) {
    double xr = 0.0, xi = 0.0;
    if (live && k < K) {
      const S2 v = src[base + tk[k]];
      xr = ldexp((double)v.x, 8 - e);
      xi = ldexp(conj ? -(double)v.y : (double)v.y, 8 - e);
    }
#pragma unroll
    for (int s = 0; s < NSL; ++s) {
      const double sr = rint(xr), si = rint(xi);
      xr = (xr - sr) * 256.0;
      xi = (xi - si) * 256.0;
      const size_t o = (((size_t)z * NSL + s) * 2) * plane + (size_t)row * Kp + k;
      out[o] = to_bits<T>(sr);
      out[o + plane] = to_bits<T>(si);
    }
  }
}

template <typename T>
struct MfmaOp;
template <>
struct MfmaOp<__bf16> {
  static __device__ __forceinline__ f4s run(u4 a, u4 b, f4s c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, a), __builtin_bit_cast(bf8, b), c, 0, 0,
                                                   0);
  }
};
template <>
struct MfmaOp<_Float16> {
  static __device__ __forceinline__ f4s run(u4 a, u4 b, f4s c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(h8, a), __builtin_bit_cast(h8, b), c, 0, 0, 0);
  }
};

struct SplitArgs {
  const unsigned short* Ap;
  const unsigned short* Bp;
  const int* ea;
  const int* eb;
  void* C;
  const int64_t* rowC;
  const int64_t* colC;
  const int64_t* bC;
  int64_t M, N, Mp, Np, Kp, batch;
  double alpha_re, alpha_im, beta_re, beta_im;
  int64_t tiles_m, tiles_n;
};

// 64 x 32 output tile per workgroup; wave w owns rows 16w..16w+15 and two
// 16-column blocks.  Operand fragments come straight from the packed planes
// (16-byte loads: 8 consecutive k of one row / column per lane).
template <typename S2, typename T, int NSL>
__global__ __launch_bounds__(256) void split_gemm_kernel(SplitArgs a) {
  constexpr int NLEV = NSL;  // levels i + j = 0 .. NSL-1 (0-based slices)
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t per = a.tiles_m * a.tiles_n;
  const int64_t tile = blockIdx.x;
  const int64_t z = tile / per, t = tile % per;
  const int64_t m0 = (t % a.tiles_m) * 64 + wave * 16, n0 = (t / a.tiles_m) * 32;
  const size_t aplane = (size_t)a.Mp * a.Kp, bplane = (size_t)a.Np * a.Kp;
  const unsigned short* Az = a.Ap + (size_t)z * NSL * 2 * aplane;
  const unsigned short* Bz = a.Bp + (size_t)z * NSL * 2 * bplane;
  const int64_t arow = m0 + (lane & 15);
  const int kl = 8 * (lane >> 4);
  double totr[2][4], toti[2][4];
#pragma unroll
  for (int cb = 0; cb < 2; ++cb)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      totr[cb][r] = 0.0;
      toti[cb][r] = 0.0;
    }
  const u4 neg = {0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
  for (int64_t kb = 0; kb < a.Kp; kb += SKB) {
    f4s accr[NLEV][2], acci[NLEV][2];
#pragma unroll
    for (int l = 0; l < NLEV; ++l)
#pragma unroll
      for (int cb = 0; cb < 2; ++cb) {
        accr[l][cb] = f4s{0.f, 0.f, 0.f, 0.f};
        acci[l][cb] = f4s{0.f, 0.f, 0.f, 0.f};
      }
    for (int ks = 0; ks < SKB; ks += SKP) {
      const int64_t k = kb + ks + kl;
      u4 ar[NSL], ai[NSL], an[NSL];
#pragma unroll
      for (int s = 0; s < NSL; ++s) {
        ar[s] = *reinterpret_cast<const u4*>(Az + (size_t)(2 * s) * aplane + (size_t)arow * a.Kp + k);
        ai[s] = *reinterpret_cast<const u4*>(Az + (size_t)(2 * s + 1) * aplane + (size_t)arow * a.Kp + k);
        an[s] = ai[s] ^ neg;  // -Im(A) slice: flip the sign bits
      }
#pragma unroll
      for (int cb = 0; cb < 2; ++cb) {
        const int64_t bcol = n0 + cb * 16 + (lane & 15);
        u4 br[NSL], bi[NSL];
#pragma unroll
        for (int s = 0; s < NSL; ++s) {
          br[s] = *reinterpret_cast<const u4*>(Bz + (size_t)(2 * s) * bplane + (size_t)bcol * a.Kp + k);
          bi[s] = *reinterpret_cast<const u4*>(Bz + (size_t)(2 * s + 1) * bplane + (size_t)bcol * a.Kp + k);
        }
        // level l = i + j - 2 (slices 0-based: i + j = l)
#pragma unroll
        for (int i = 0; i < NSL; ++i)
#pragma unroll
          for (int j = 0; j < NSL; ++j) {
            if (i + j >= NSL) continue;
            const int l = i + j;
            accr[l][cb] = MfmaOp<T>::run(ar[i], br[j], accr[l][cb]);
            accr[l][cb] = MfmaOp<T>::run(an[i], bi[j], accr[l][cb]);
            acci[l][cb] = MfmaOp<T>::run(ar[i], bi[j], acci[l][cb]);
            acci[l][cb] = MfmaOp<T>::run(ai[i], br[j], acci[l][cb]);
          }
      }
    }
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        double sr = 0.0, si = 0.0, w = 1.0;
#pragma unroll
        for (int l = 0; l < NLEV; ++l) {
          sr += w * (double)accr[l][cb][r];
          si += w * (double)acci[l][cb][r];
          w *= 0x1p-8;
        }
        totr[cb][r] += sr;
        toti[cb][r] += si;
      }
  }
  // epilogue: D layout rows 4*(lane/16) + r, column lane % 16
  S2* C = static_cast<S2*>(a.C);
  const int64_t offCz = a.bC[z];
  const bool use_beta = (a.beta_re != 0.0 || a.beta_im != 0.0);
#pragma unroll
  for (int cb = 0; cb < 2; ++cb) {
    const int64_t col = n0 + cb * 16 + (lane & 15);
    if (col >= a.N) continue;
    const int fcol = a.eb[z * a.Np + col];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t row = m0 + 4 * (lane >> 4) + r;
      if (row >= a.M) continue;
      const int sc = a.ea[z * a.Mp + row] + fcol - 16;
      const double cr = ldexp(totr[cb][r], sc), ci = ldexp(toti[cb][r], sc);
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

template <typename S2, typename T, int NSL>
static int run_split(const GemmLaunch& g, hipStream_t s, void* ws) {
  const SplitLayout L = split_layout(g);
  char* w = static_cast<char*>(ws);
  unsigned short* Ap = reinterpret_cast<unsigned short*>(w);
  w += align256(L.a_bytes);
  unsigned short* Bp = reinterpret_cast<unsigned short*>(w);
  w += align256(L.b_bytes);
  int* ea = reinterpret_cast<int*>(w);
  w += align256(L.ea_bytes);
  int* eb = reinterpret_cast<int*>(w);
  {
    const int64_t rows = L.Mp * g.batch;
    hipLaunchKernelGGL((split_pack_kernel<S2, T, NSL>), dim3((unsigned)((rows + 3) / 4)), dim3(256), 0, s,
                       static_cast<const S2*>(g.A), g.t.rowA, g.t.kA, g.t.bA, g.M, g.K, L.Mp, L.Kp, g.batch, g.conjA,
                       Ap, ea);
    TENET_HIP_CHECK(hipGetLastError());
  }
  {
    const int64_t rows = L.Np * g.batch;
    hipLaunchKernelGGL((split_pack_kernel<S2, T, NSL>), dim3((unsigned)((rows + 3) / 4)), dim3(256), 0, s,
                       static_cast<const S2*>(g.B), g.t.colB, g.t.kB, g.t.bB, g.N, g.K, L.Np, L.Kp, g.batch, g.conjB,
                       Bp, eb);
    TENET_HIP_CHECK(hipGetLastError());
  }
  SplitArgs a;
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
  a.Mp = L.Mp;
  a.Np = L.Np;
  a.Kp = L.Kp;
  a.batch = g.batch;
  a.alpha_re = g.alpha_re;
  a.alpha_im = g.alpha_im;
  a.beta_re = g.beta_re;
  a.beta_im = g.beta_im;
  a.tiles_m = L.Mp / 64;
  a.tiles_n = L.Np / 32;
  const int64_t total = a.tiles_m * a.tiles_n * g.batch;
  if (total > 0x7fffffffLL) {
    set_error("split contraction too large for one launch");
    return TENET_ERR_UNSUPPORTED;
  }
  hipLaunchKernelGGL((split_gemm_kernel<S2, T, NSL>), dim3((unsigned)total), dim3(256), 0, s, a);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

int launch_split_gemm(const GemmLaunch& g, int dtype, int prec, hipStream_t s, void* ws, size_t ws_bytes) {
  if (!ws || ws_bytes < gemm_workspace_bytes(g, dtype, prec)) {
    set_error("split contraction: workspace too small");
    return TENET_ERR_OOM;
  }
  if (g.K == 0) {
    set_error("split contraction with an empty contracted index");
    return TENET_ERR_UNSUPPORTED;
  }
  if (dtype == TENET_C128)
    return prec == TENET_PREC_SPLIT_BF16 ? run_split<double2, __bf16, 6>(g, s, ws)
                                         : run_split<double2, _Float16, 6>(g, s, ws);
  return prec == TENET_PREC_SPLIT_BF16 ? run_split<float2, __bf16, 3>(g, s, ws) : run_split<float2, _Float16, 3>(g, s, ws);
}

}  // namespace tenet
