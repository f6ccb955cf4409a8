// K2: thin Householder QR in FP64 (complex), the factorisation behind
// Tenet's canonize! moves (Muscle.tensor_qr_thin, canonize.jl:58).
//
// Reflectors follow Golub & Van Loan (Matrix Computations, 4th ed., 5.1.2)
// in the LAPACK convention: H = I - tau v v^H with v(0) = 1, beta real,
// beta = -sign(Re alpha) * ||(alpha, x)||, applied as H^H from the left.
// Q = H_1 H_2 ... H_k [I; 0] is accumulated backwards.
//
// One workgroup owns one matrix (batched over blockIdx.x).  The matrix
// lives in a FP64 work copy in global memory; the current reflector is
// staged in LDS and each wave owns a set of trailing columns, so the dot
// products need only wave-level (64-lane) reductions.
#include <hip/hip_ext.h>

#include "../../include/tenet_hip.h"
#include "internal.h"

#include <utility>

namespace tenet {

constexpr int QR_THREADS = 256;
constexpr int QR_MAX_M = 4096;  // LDS holds one reflector of up to this length

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

struct Cd {
  double r, i;
};

__device__ __forceinline__ void ld2(const double* p, double& r, double& i) {
  const double2 v = *reinterpret_cast<const double2*>(p);
  r = v.x;
  i = v.y;
}
__device__ __forceinline__ void st2(double* p, double r, double i) {
  double2 v;
  v.x = r;
  v.y = i;
  *reinterpret_cast<double2*>(p) = v;
}

// In-place Householder factorisation of W (m x n complex, ld = m).
__global__ __launch_bounds__(QR_THREADS) void qr_factor_kernel(double* __restrict__ Wall, double* __restrict__ tauall,
                                                                int m, int n) {
  __shared__ double vr[QR_MAX_M], vi[QR_MAX_M];
  __shared__ double red[QR_THREADS / 64];
  __shared__ double s_tau_r, s_tau_i;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int k = min(m, n);
  double* W = Wall + (size_t)blockIdx.x * (size_t)m * n * 2;
  double* tau = tauall + (size_t)blockIdx.x * k * 2;

  for (int j = 0; j < k; ++j) {
    // ||x||^2 of the sub-diagonal part of column j
    double part = 0.0;
    for (int i = j + 1 + tid; i < m; i += QR_THREADS) {
      double xr, xi;
      ld2(W + 2 * ((size_t)j * m + i), xr, xi);
      part += xr * xr + xi * xi;
    }
    part = wave_sum(part);
    if (lane == 0) red[wave] = part;
    __syncthreads();
    if (tid == 0) {
      double xn2 = 0.0;
      for (int w = 0; w < QR_THREADS / 64; ++w) xn2 += red[w];
      double ar, ai;
      ld2(W + 2 * ((size_t)j * m + j), ar, ai);
      double tr = 0.0, ti = 0.0, sr = 0.0, si = 0.0, beta = ar;
      if (xn2 == 0.0 && ai == 0.0) {
        tr = 0.0;
        ti = 0.0;  // H = I
      } else {
        const double nrm = sqrt(ar * ar + ai * ai + xn2);
        beta = (ar >= 0.0) ? -nrm : nrm;
        tr = (beta - ar) / beta;
        ti = -ai / beta;
        // scale = 1 / (alpha - beta)
        const double dr = ar - beta, di = ai;
        const double den = dr * dr + di * di;
        sr = dr / den;
        si = -di / den;
      }
      s_tau_r = tr;
      s_tau_i = ti;
      red[0] = sr;
      red[1] = si;
      red[2] = beta;
      st2(tau + 2 * j, tr, ti);
    }
    __syncthreads();
    const double sr = red[0], si = red[1], beta = red[2];
    const double tr = s_tau_r, ti = s_tau_i;
    const bool identity = (tr == 0.0 && ti == 0.0);
    // v = [1; x * scale], stored below the diagonal of W and staged in LDS
    for (int i = j + tid; i < m; i += QR_THREADS) {
      if (i == j) {
        vr[i] = 1.0;
        vi[i] = 0.0;
        if (!identity) st2(W + 2 * ((size_t)j * m + j), beta, 0.0);
      } else {
        double xr, xi;
        ld2(W + 2 * ((size_t)j * m + i), xr, xi);
        double nr = xr, ni = xi;
        if (!identity) {
          nr = xr * sr - xi * si;
          ni = xr * si + xi * sr;
          st2(W + 2 * ((size_t)j * m + i), nr, ni);
        }
        vr[i] = nr;
        vi[i] = ni;
      }
    }
    __syncthreads();
    if (!identity) {
      // A(j:, c) -= conj(tau) * v * (v^H A(j:, c)), one wave per column
      for (int c = j + 1 + wave; c < n; c += QR_THREADS / 64) {
        double* col = W + 2 * (size_t)c * m;
        double dr = 0.0, di = 0.0;
        for (int i = j + lane; i < m; i += 64) {
          double xr, xi;
          ld2(col + 2 * i, xr, xi);
          // conj(v_i) * x_i
          dr += vr[i] * xr + vi[i] * xi;
          di += vr[i] * xi - vi[i] * xr;
        }
        dr = wave_sum(dr);
        di = wave_sum(di);
        // f = conj(tau) * dot
        const double fr = tr * dr + ti * di;
        const double fi = tr * di - ti * dr;
        for (int i = j + lane; i < m; i += 64) {
          double xr, xi;
          ld2(col + 2 * i, xr, xi);
          xr -= vr[i] * fr - vi[i] * fi;
          xi -= vr[i] * fi + vi[i] * fr;
          st2(col + 2 * i, xr, xi);
        }
      }
    }
    __syncthreads();
  }
}

template <typename S2>
__device__ __forceinline__ void store_c(void* base, size_t idx, double r, double i) {
  S2 v;
  v.x = r;
  v.y = i;
  static_cast<S2*>(base)[idx] = v;
}

// R (k x n) from the upper triangle; Q (m x k) = H_1 ... H_k [I; 0].
// Q is formed in place in a scratch m x k region that follows W.
template <typename S2>
__global__ __launch_bounds__(QR_THREADS) void qr_extract_kernel(const double* __restrict__ Wall,
                                                                 const double* __restrict__ tauall, int m, int n,
                                                                 double* __restrict__ Qscr, void* Qout, int64_t ldq,
                                                                 void* Rout, int64_t ldr, int flags) {
  __shared__ double vr[QR_MAX_M], vi[QR_MAX_M];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int k = min(m, n);
  const size_t b = blockIdx.x;
  const double* W = Wall + b * (size_t)m * n * 2;
  const double* tau = tauall + b * k * 2;
  double* Q = Qscr + b * (size_t)m * k * 2;
  S2* Qo = static_cast<S2*>(Qout) + b * (size_t)ldq * ((flags & TENET_QR_Q_ADJ) ? m : k);
  S2* Ro = static_cast<S2*>(Rout) + b * (size_t)ldr * ((flags & TENET_QR_R_ADJ) ? k : n);

  const bool radj = (flags & TENET_QR_R_ADJ) != 0, qadj = (flags & TENET_QR_Q_ADJ) != 0;
  for (int c = 0; c < n; ++c)
    for (int r = tid; r < k; r += QR_THREADS) {
      double xr = 0.0, xi = 0.0;
      if (r <= c) ld2(W + 2 * ((size_t)c * m + r), xr, xi);
      S2 v;
      v.x = xr;
      v.y = radj ? -xi : xi;
      if (radj)
        Ro[(size_t)r * ldr + c] = v;  // R^H is n x k
      else
        Ro[(size_t)c * ldr + r] = v;
    }
  // Q <- [I; 0]
  for (size_t e = tid; e < (size_t)m * k; e += QR_THREADS) {
    const int r = (int)(e % m), c = (int)(e / m);
    st2(Q + 2 * e, r == c ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  for (int j = k - 1; j >= 0; --j) {
    const double tr = tau[2 * j], ti = tau[2 * j + 1];
    if (tr == 0.0 && ti == 0.0) continue;
    for (int i = j + tid; i < m; i += QR_THREADS) {
      if (i == j) {
        vr[i] = 1.0;
        vi[i] = 0.0;
      } else {
        ld2(W + 2 * ((size_t)j * m + i), vr[i], vi[i]);
      }
    }
    __syncthreads();
    for (int c = j + wave; c < k; c += QR_THREADS / 64) {
      double* col = Q + 2 * (size_t)c * m;
      double dr = 0.0, di = 0.0;
      for (int i = j + lane; i < m; i += 64) {
        double xr, xi;
        ld2(col + 2 * i, xr, xi);
        dr += vr[i] * xr + vi[i] * xi;
        di += vr[i] * xi - vi[i] * xr;
      }
      dr = wave_sum(dr);
      di = wave_sum(di);
      const double fr = tr * dr - ti * di;  // tau * dot
      const double fi = tr * di + ti * dr;
      for (int i = j + lane; i < m; i += 64) {
        double xr, xi;
        ld2(col + 2 * i, xr, xi);
        xr -= vr[i] * fr - vi[i] * fi;
        xi -= vr[i] * fi + vi[i] * fr;
        st2(col + 2 * i, xr, xi);
      }
    }
    __syncthreads();
  }
  for (size_t e = tid; e < (size_t)m * k; e += QR_THREADS) {
    const int r = (int)(e % m), c = (int)(e / m);
    double xr, xi;
    ld2(Q + 2 * e, xr, xi);
    S2 v;
    v.x = xr;
    v.y = qadj ? -xi : xi;
    if (qadj)
      Qo[(size_t)r * ldq + c] = v;  // Q^H is k x m
    else
      Qo[(size_t)c * ldq + r] = v;
  }
}

int launch_qr_factor(double* W, double* tau, int64_t m, int64_t n, int64_t batch, hipStream_t s) {
  if (m > QR_MAX_M) {
    set_error("tenet_qr_thin: m exceeds the single-workgroup limit (4096)");
    return TENET_ERR_UNSUPPORTED;
  }
  if (batch <= 0 || m <= 0 || n <= 0) return TENET_OK;
  hipLaunchKernelGGL(qr_factor_kernel, dim3((unsigned)batch), dim3(QR_THREADS), 0, s, W, tau, (int)m, (int)n);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

// W region must be followed by m*k*batch complex doubles of scratch for Q.
int launch_qr_extract(const double* W, const double* tau, int64_t m, int64_t n, int64_t batch, void* Q,
                      int64_t ldq, void* R, int64_t ldr, int dtype, int flags, hipStream_t s) {
  if (batch <= 0 || m <= 0 || n <= 0) return TENET_OK;
  double* Qscr = const_cast<double*>(W) + (size_t)batch * m * n * 2;
  if (dtype == TENET_C128)
    hipLaunchKernelGGL(qr_extract_kernel<double2>, dim3((unsigned)batch), dim3(QR_THREADS), 0, s, W, tau, (int)m,
                       (int)n, Qscr, Q, ldq, R, ldr, flags);
  else
    hipLaunchKernelGGL(qr_extract_kernel<float2>, dim3((unsigned)batch), dim3(QR_THREADS), 0, s, W, tau, (int)m,
                       (int)n, Qscr, Q, ldq, R, ldr, flags);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

}  // namespace tenet

// ---------------------------------------------------------------------------
// Blocked Householder QR (compact WY, Golub & Van Loan 5.2.3; LAPACK zlarft
// conventions): panels of QB columns are factored by one workgroup with the
// panel resident in registers, and the trailing matrix and Q are updated by
// many workgroups, one 16-column block each:
//   panel:    H_1 ... H_jb = I - V T V^H      (T upper triangular)
//   trailing: A <- (I - V T V^H)^H A = A - V (T^H (V^H A))
//   Q build:  Q <- (I - V T V^H) Q, panels in reverse order, Q = [I; 0]
// ---------------------------------------------------------------------------
namespace tenet {

constexpr int QB = 16;            // panel width
constexpr int QP_MAX_ROWS = 1024;  // panel rows held in registers (up to four per thread of 256)

// Panel v3: 256 threads, thread t owns panel rows t and t + 256 in
// registers (up to 512 rows x QB columns), the column loop fully unrolled so
// every register index is a compile-time constant and the trailing work is
// triangular.  One block reduction per column: with x the column below the
// diagonal and P the pivot row (written to LDS by its owner), every dot the
// column needs follows from e_c = x^H a_c (rows below the diagonal):
//   ||x||^2 = e_jj,  v^H a_c = conj(s) e_c + P_c  (v = s x below, 1 on the
//   diagonal),  v_c^H v = conj(P_c) + s conj(e_c)  (c < jj, for T).
// The 32 real values of e are transpose-reduced inside each wave (levels
// pairing lanes l ^ 15, l ^ 7, l ^ 2, l ^ 1 by DPP, l ^ 16 and l ^ 32 by
// permlane swaps; lane l ends with the wave total of value 16 b3 + 8 b2 + 4 b1 +
// 2 b0 + b4, b_i the bits of l), the four wave totals meet in LDS (double
// buffered by column parity): one barrier per column.
template <int CTRL>
__device__ __forceinline__ double dpp_d(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Lane exchanges on the VALU (gfx950 permlane swaps of the 16-lane rows and
// the 32-lane halves).  xpair16_d(a, b): lanes of even rows get a(l) +
// a(l ^ 16), lanes of odd rows b(l) + b(l ^ 16); xsum16_d / xsum32_d:
// x(l) + x(l ^ 16) / x(l) + x(l ^ 32).  Both lanes of a pair add the same two
// values in the same order.
__device__ __forceinline__ double xpair16_d(double a, double b) {
  const long long ba = __double_as_longlong(a), bb = __double_as_longlong(b);
  const auto pl = __builtin_amdgcn_permlane16_swap((unsigned)(ba & 0xffffffffll), (unsigned)(bb & 0xffffffffll),
                                                   false, false);
  const auto ph = __builtin_amdgcn_permlane16_swap((unsigned)(ba >> 32), (unsigned)(bb >> 32), false, false);
  const double e = __longlong_as_double(((long long)ph[0] << 32) | (long long)pl[0]);
  const double o = __longlong_as_double(((long long)ph[1] << 32) | (long long)pl[1]);
  return e + o;
}

__device__ __forceinline__ double xsum16_d(double x) { return xpair16_d(x, x); }

__device__ __forceinline__ double xsum32_d(double x) {
  const long long b = __double_as_longlong(x);
  const unsigned lo = (unsigned)(b & 0xffffffffll), hi = (unsigned)(b >> 32);
  const auto pl = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
  const auto ph = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
  const double e = __longlong_as_double(((long long)ph[0] << 32) | (long long)pl[0]);
  const double o = __longlong_as_double(((long long)ph[1] << 32) | (long long)pl[1]);
  return e + o;
}

__device__ __forceinline__ double readlane_d(double x, int src) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), src);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Sum over the 64 lanes, the same value in every lane: quads and the two
// halves of each 16-lane row by DPP, then the rows and halves of the wave.
__device__ __forceinline__ double wsum64(double v) {
  v += dpp_d<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_d<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_d<0x141>(v);  // row_half_mirror: the other quad of the 8
  v += dpp_d<0x140>(v);  // row_mirror: the other 8 of the 16
  return xsum32_d(xsum16_d(v));
}

template <int H, int CTRL, int BIT>
__device__ __forceinline__ void halve_dpp(double (&v)[2 * QB], int lane) {
  const bool hi = (lane & BIT) != 0;
#pragma unroll
  for (int i = 0; i < H; ++i) {
    const double keep = hi ? v[H + i] : v[i];
    const double send = hi ? v[i] : v[H + i];
    v[i] = keep + dpp_d<CTRL>(send);
  }
}

__device__ __forceinline__ void wave_transpose_reduce32(double (&v)[2 * QB], int lane) {
  static_assert(2 * QB == 32, "the network below reduces 32 values");
  halve_dpp<16, 0x140, 8>(v, lane);  // row_mirror: l ^ 15
  halve_dpp<8, 0x141, 4>(v, lane);   // row_half_mirror: l ^ 7
  halve_dpp<4, 0x4E, 2>(v, lane);    // quad_perm [2,3,0,1]: l ^ 2
  halve_dpp<2, 0xB1, 1>(v, lane);    // quad_perm [1,0,3,2]: l ^ 1
  v[0] = xsum32_d(xpair16_d(v[0], v[1]));  // l ^ 16 (keeping v[0] on even rows, v[1] on odd), then l ^ 32
}

constexpr int QP3_THREADS = 256;

// One column of the v3 panel (JJ compile-time: static register indices;
// thread t holds panel rows t + 256 r, r < R).
template <int JJ, int R>
__device__ __forceinline__ void qr_panel_v3_step(double (&ar)[R][QB], double (&ai)[R][QB],
                                                 double (*part)[QP3_THREADS / 64][2 * QB],
                                                 double (*prow)[2 * QB], double (*Tr_s)[QB], double (*Ti_s)[QB],
                                                 double* __restrict__ tau, int tid, int lane, int wave, int idx, const int (&rowv)[R],
                                                 int j0) {
  constexpr int jj = JJ;
  const int par = jj & 1;
  if (tid == jj) {  // pivot row jj is row slot 0 of thread jj
#pragma unroll
    for (int c = 0; c < QB; ++c) {
      prow[par][2 * c] = ar[0][c];
      prow[par][2 * c + 1] = ai[0][c];
    }
  }
  // e_c = sum over this thread's rows below the diagonal of conj(x_r) a_c[r]
  double e[2 * QB];
#pragma unroll
  for (int c = 0; c < QB; ++c) e[2 * c] = e[2 * c + 1] = 0.0;
#pragma unroll
  for (int sl = 0; sl < R; ++sl) {
    const int row = rowv[sl];
    if (row > jj) {
      const double xr = ar[sl][jj], xi = ai[sl][jj];
#pragma unroll
      for (int c = 0; c < QB; ++c) {
        e[2 * c] += xr * ar[sl][c] + xi * ai[sl][c];
        e[2 * c + 1] += xr * ai[sl][c] - xi * ar[sl][c];
      }
    }
  }
  wave_transpose_reduce32(e, lane);
  if (lane < 32) part[par][wave][idx] = e[0];
  __syncthreads();
  // column totals (fixed wave order) and the pivot row
  auto E = [&](int i) {
    return ((part[par][0][i] + part[par][1][i]) + part[par][2][i]) + part[par][3][i];
  };
  const double alr = prow[par][2 * jj], ali = prow[par][2 * jj + 1];
  const double xn2 = E(2 * jj);
  double t_r = 0.0, t_i = 0.0, sr = 1.0, si = 0.0, beta = alr;
  const bool ident = (xn2 == 0.0 && ali == 0.0);
  if (!ident) {
    const double nrm = sqrt(alr * alr + ali * ali + xn2);
    beta = (alr >= 0.0) ? -nrm : nrm;
    t_r = (beta - alr) / beta;
    t_i = -ali / beta;
    const double dr = alr - beta, di = ali;
    const double den = dr * dr + di * di;
    sr = dr / den;
    si = -di / den;
  }
  // trailing columns: w_c = v^H a_c = conj(s) e_c + P_c, f_c = conj(tau) w_c,
  // a_c[r] -= v_r f_c on rows r >= jj
  if (!ident) {
#pragma unroll
    for (int c = jj + 1; c < QB; ++c) {
      const double er = E(2 * c), ei = E(2 * c + 1);
      const double wr = sr * er + si * ei + prow[par][2 * c];
      const double wi = sr * ei - si * er + prow[par][2 * c + 1];
      const double fr = t_r * wr + t_i * wi;
      const double fi = t_r * wi - t_i * wr;
#pragma unroll
      for (int sl = 0; sl < R; ++sl) {
        const int row = rowv[sl];
        if (row == jj) {
          ar[sl][c] -= fr;
          ai[sl][c] -= fi;
        } else if (row > jj) {
          const double xr = ar[sl][jj], xi = ai[sl][jj];
          const double vr = xr * sr - xi * si, vi = xr * si + xi * sr;
          ar[sl][c] -= vr * fr - vi * fi;
          ai[sl][c] -= vr * fi + vi * fr;
        }
      }
    }
  }
  // the reflector into column jj (below the diagonal), beta on it
#pragma unroll
  for (int sl = 0; sl < R; ++sl) {
    const int row = rowv[sl];
    if (row > jj && !ident) {
      const double xr = ar[sl][jj], xi = ai[sl][jj];
      ar[sl][jj] = xr * sr - xi * si;
      ai[sl][jj] = xr * si + xi * sr;
    } else if (row == jj && !ident) {
      ar[sl][jj] = beta;
      ai[sl][jj] = 0.0;
    }
  }
  // T[a][jj] = -tau sum_{c=a}^{jj-1} T[a][c] (v_c^H v), v_c^H v = conj(P_c) + s conj(e_c)
  if (tid < jj) {
    double wr = 0.0, wi = 0.0;
#pragma unroll
    for (int c = 0; c < jj; ++c) {
      if (c >= tid) {
        const double er = E(2 * c), ei = E(2 * c + 1);
        const double dr = prow[par][2 * c] + (sr * er + si * ei);
        const double di = -prow[par][2 * c + 1] + (si * er - sr * ei);
        wr += Tr_s[tid][c] * dr - Ti_s[tid][c] * di;
        wi += Tr_s[tid][c] * di + Ti_s[tid][c] * dr;
      }
    }
    Tr_s[tid][jj] = -(t_r * wr - t_i * wi);
    Ti_s[tid][jj] = -(t_r * wi + t_i * wr);
  }
  if (tid == jj) {
    Tr_s[jj][jj] = t_r;
    Ti_s[jj][jj] = t_i;
    tau[2 * (j0 + jj)] = t_r;
    tau[2 * (j0 + jj) + 1] = t_i;
  }
  }

template <int R, int... J>
__device__ __forceinline__ void qr_panel_v3_cols(std::integer_sequence<int, J...>, int jb, double (&ar)[R][QB],
                                                 double (&ai)[R][QB], double (*part)[QP3_THREADS / 64][2 * QB],
                                                 double (*prow)[2 * QB], double (*Tr_s)[QB], double (*Ti_s)[QB],
                                                 double* __restrict__ tau, int tid, int lane, int wave, int idx,
                                                 const int (&rowv)[R], int j0) {
  ((J < jb ? qr_panel_v3_step<J, R>(ar, ai, part, prow, Tr_s, Ti_s, tau, tid, lane, wave, idx, rowv, j0) : void()),
   ...);
}

template <int R>
__global__ __launch_bounds__(QP3_THREADS) void qr_panel_v3_kernel(double* __restrict__ W, double* __restrict__ tau,
                                                                   double* __restrict__ Tout, int m, int j0, int jb) {
  __shared__ double part[2][QP3_THREADS / 64][2 * QB];
  __shared__ double prow[2][2 * QB];
  __shared__ double Tr_s[QB][QB], Ti_s[QB][QB];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int mr = m - j0;
  double2* Wc = reinterpret_cast<double2*>(W);
  double ar[R][QB], ai[R][QB];
#pragma unroll
  for (int c = 0; c < QB; ++c)
#pragma unroll
    for (int sl = 0; sl < R; ++sl) {
      const int row = tid + sl * QP3_THREADS;
      ar[sl][c] = ai[sl][c] = 0.0;
      if (c < jb && row < mr) {
        const double2 v = Wc[(size_t)(j0 + c) * m + j0 + row];
        ar[sl][c] = v.x;
        ai[sl][c] = v.y;
      }
    }
  for (int e = tid; e < QB * QB; e += QP3_THREADS) {
    Tr_s[e % QB][e / QB] = 0.0;
    Ti_s[e % QB][e / QB] = 0.0;
  }
  const int idx = 16 * ((lane >> 3) & 1) + 8 * ((lane >> 2) & 1) + 4 * ((lane >> 1) & 1) + 2 * (lane & 1) +
                  ((lane >> 4) & 1);
  int rowv[R];
#pragma unroll
  for (int sl = 0; sl < R; ++sl) rowv[sl] = tid + sl * QP3_THREADS;
  qr_panel_v3_cols<R>(std::make_integer_sequence<int, QB>{}, jb, ar, ai, part, prow, Tr_s, Ti_s, tau, tid, lane, wave,
                      idx, rowv, j0);
  __syncthreads();
#pragma unroll
  for (int c = 0; c < QB; ++c)
#pragma unroll
    for (int sl = 0; sl < R; ++sl) {
      const int row = tid + sl * QP3_THREADS;
      if (c < jb && row < mr) {
        double2 v;
        v.x = ar[sl][c];
        v.y = ai[sl][c];
        Wc[(size_t)(j0 + c) * m + j0 + row] = v;
      }
    }
  for (int e = tid; e < QB * QB; e += QP3_THREADS) {
    const int i = e % QB, j = e / QB;
    Tout[2 * e] = (i < jb && j < jb) ? Tr_s[i][j] : 0.0;
    Tout[2 * e + 1] = (i < jb && j < jb) ? Ti_s[i][j] : 0.0;
  }
}


// Panel v4 (panel rows <= 512): one wave per panel column (QB waves), lane l
// of wave c holding rows l + 64 r (r < RL, panel-relative) of column c in
// registers.  Step j: every wave right of j applies reflector j (v, tau from
// LDS) to its column, with a 64-lane sum for v^H a_c; the wave of column
// j + 1 then forms reflector j + 1 and publishes it; the waves left of j form
// w_c = v_c^H v_j (their registers hold v_c below the diagonal) for T's
// column j, which the then idle wave j + 1 assembles one step later
// (T(a, j) = -tau_j sum_{c=a}^{j-1} T(a, c) w_c, T(j, j) = tau_j: zlarft's
// forward columnwise rule).  One barrier per column; the outputs are those of
// qr_panel_v3_kernel.
template <int RL>
__device__ __forceinline__ void qp4_reflector(double (&ar)[RL], double (&ai)[RL], int jj, int lane,
                                              double2 (*Vs)[64 * RL], double2* taus, double* __restrict__ tau,
                                              int j0) {
  double part = 0.0;
#pragma unroll
  for (int r = 0; r < RL; ++r) {
    const double t = fma(ar[r], ar[r], ai[r] * ai[r]);
    part += (lane + 64 * r > jj) ? t : 0.0;
  }
  const double xn = readlane_d(wsum64(part), 0);
  const double alr = readlane_d(ar[0], jj), ali = readlane_d(ai[0], jj);  // row jj < 64: lane jj, slot 0
  double tr = 0.0, ti = 0.0, scr = 0.0, sci = 0.0, beta = alr;
  if (!(xn == 0.0 && ali == 0.0)) {
    const double nrm = sqrt(fma(alr, alr, fma(ali, ali, xn)));
    beta = (alr >= 0.0) ? -nrm : nrm;
    const double dr = alr - beta, di = ali;
    const double ib = 1.0 / beta, iden = 1.0 / fma(dr, dr, di * di);
    tr = (beta - alr) * ib;
    ti = -ali * ib;
    scr = dr * iden;  // 1 / (alpha - beta)
    sci = -di * iden;
  }
#pragma unroll
  for (int r = 0; r < RL; ++r) {
    const int row = lane + 64 * r;
    const bool below = row > jj, diag = row == jj;
    double2 v;
    v.x = below ? fma(ar[r], scr, -ai[r] * sci) : (diag ? 1.0 : 0.0);
    v.y = below ? fma(ar[r], sci, ai[r] * scr) : 0.0;
    ar[r] = below ? v.x : (diag ? beta : ar[r]);
    ai[r] = below ? v.y : (diag ? 0.0 : ai[r]);
    Vs[jj & 1][row] = v;  // double buffered: waves may still read reflector jj - 1
  }
  if (lane == 0) {
    double2 t;
    t.x = tr;
    t.y = ti;
    taus[jj] = t;
    tau[2 * (j0 + jj)] = tr;
    tau[2 * (j0 + jj) + 1] = ti;
  }
}

// T column jj from the w_c = v_c^H v_jj (lanes a <= jj of one wave)
__device__ __forceinline__ void qp4_tcol(int jj, int lane, const double2* taus, const double2* w,
                                         double2 (*Ts)[QB]) {
  if (lane > jj) return;
  const double2 t = taus[jj];
  if (lane == jj) {
    Ts[jj][jj] = t;
    return;
  }
  double sr = 0.0, si = 0.0;
  for (int c = lane; c < jj; ++c) {  // T(a, c) w_c, a = lane
    const double2 x = Ts[c][lane], y = w[c];
    sr = fma(x.x, y.x, fma(-x.y, y.y, sr));
    si = fma(x.x, y.y, fma(x.y, y.x, si));
  }
  double2 o;
  o.x = -(t.x * sr - t.y * si);
  o.y = -(t.x * si + t.y * sr);
  Ts[jj][lane] = o;
}

template <int RL>
__device__ __forceinline__ void qp4_panel(double* __restrict__ W, double* __restrict__ tau, double* __restrict__ Tout,
                                          int m, int j0, int jb) {
  __shared__ double2 Vs[2][64 * RL];  // reflector j in Vs[j & 1], panel rows
  __shared__ double2 taus[QB];
  __shared__ double2 Ws[2][QB];    // w_c = v_c^H v_j by the parity of j
  __shared__ double2 Ts[QB][QB];   // Ts[c][a] = T(a, c)
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // panel column
  const int mr = m - j0;
  double2* Wc = reinterpret_cast<double2*>(W);
  double ar[RL], ai[RL];
#pragma unroll
  for (int r = 0; r < RL; ++r) {  // branch-free: lanes outside the panel read element 0 and drop it
    const int row = lane + 64 * r;
    const bool ok = wave < jb && row < mr;
    const double2 v = Wc[ok ? (size_t)(j0 + wave) * m + j0 + row : 0];
    ar[r] = ok ? v.x : 0.0;
    ai[r] = ok ? v.y : 0.0;
  }
  for (int e = tid; e < QB * QB; e += 64 * QB) {
    double2 z;
    z.x = 0.0;
    z.y = 0.0;
    Ts[e / QB][e % QB] = z;
  }
  if (wave == 0) qp4_reflector<RL>(ar, ai, 0, lane, Vs, taus, tau, j0);
  __syncthreads();
  for (int j = 0; j < jb; ++j) {
    const double2 t = taus[j];
    const double2* Vj = Vs[j & 1];
    double vr[RL], vi[RL];
#pragma unroll
    for (int r = 0; r < RL; ++r) {
      const double2 v = Vj[lane + 64 * r];
      vr[r] = v.x;
      vi[r] = v.y;
    }
    double r0 = 0.0, i0 = 0.0;  // v_j^H a (trailing) or v_c^H v_j (left of j)
    if (wave > j && wave < jb) {
#pragma unroll
      for (int r = 0; r < RL; ++r) {
        r0 = fma(vr[r], ar[r], fma(vi[r], ai[r], r0));
        i0 = fma(vr[r], ai[r], fma(-vi[r], ar[r], i0));
      }
      const double dr = wsum64(r0), di = wsum64(i0);
      const double fr = t.x * dr + t.y * di, fi = t.x * di - t.y * dr;  // conj(tau) v^H a
#pragma unroll
      for (int r = 0; r < RL; ++r) {
        ar[r] = fma(-vr[r], fr, fma(vi[r], fi, ar[r]));
        ai[r] = fma(-vr[r], fi, fma(-vi[r], fr, ai[r]));
      }
      if (wave == j + 1) qp4_reflector<RL>(ar, ai, j + 1, lane, Vs, taus, tau, j0);
    } else if (wave < j) {
      // v_c is 0 above row c and v_j above row j > c: the registers' R part meets zeros
#pragma unroll
      for (int r = 0; r < RL; ++r) {
        r0 = fma(ar[r], vr[r], fma(ai[r], vi[r], r0));
        i0 = fma(ar[r], vi[r], fma(-ai[r], vr[r], i0));
      }
      const double wr = wsum64(r0), wi = wsum64(i0);
      if (lane == 0) {
        double2 w;
        w.x = wr;
        w.y = wi;
        Ws[j & 1][wave] = w;
      }
    } else if (wave == j && j > 0) {
      qp4_tcol(j - 1, lane, taus, Ws[(j - 1) & 1], Ts);
    }
    __syncthreads();
  }
  if (wave == 0 && jb > 0) qp4_tcol(jb - 1, lane, taus, Ws[(jb - 1) & 1], Ts);
  __syncthreads();
  if (wave < jb) {
#pragma unroll
    for (int r = 0; r < RL; ++r) {
      const int row = lane + 64 * r;
      if (row < mr) {
        double2 v;
        v.x = ar[r];
        v.y = ai[r];
        Wc[(size_t)(j0 + wave) * m + j0 + row] = v;
      }
    }
  }
  for (int e = tid; e < QB * QB; e += 64 * QB) {
    const int i = e % QB, jc = e / QB;
    const double2 x = Ts[jc][i];
    Tout[2 * e] = (i < jb && jc < jb) ? x.x : 0.0;
    Tout[2 * e + 1] = (i < jb && jc < jb) ? x.y : 0.0;
  }
}

template <int RL>
__global__ __launch_bounds__(64 * QB) void qr_panel_v4_kernel(double* __restrict__ W, double* __restrict__ tau,
                                                             double* __restrict__ Tout, int m, int j0, int jb) {
  qp4_panel<RL>(W, tau, Tout, m, j0, jb);
}

// Panel p of every problem of a batch (one workgroup per problem; problems
// with fewer than p + 1 panels return at once)
template <int RL>
__global__ __launch_bounds__(64 * QB) void qr_panel_v4_batched_kernel(const QrBatchProb* __restrict__ probs, int p) {
  const QrBatchProb& q = probs[blockIdx.x];
  const int k = min(q.m, q.n), j0 = p * QB;
  if (j0 >= k) return;
  qp4_panel<RL>(q.W, q.tau, q.T + (size_t)p * QB * QB * 2, q.m, j0, min(QB, k - j0));
}

// Q (m x k) = [I; 0] in FP64
__global__ void qr_eye_kernel(double* __restrict__ Q, int m, int k) {
  const size_t total = (size_t)m * k;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const int r = (int)(e % m), c = (int)(e / m);
    double2 v;
    v.x = (r == c) ? 1.0 : 0.0;
    v.y = 0.0;
    reinterpret_cast<double2*>(Q)[e] = v;
  }
}

// Copy Q (FP64, m x k) and R (upper triangle of W) to the outputs
template <typename S2>
__global__ void qr_store_kernel(const double* __restrict__ W, const double* __restrict__ Q, int m, int n, void* Qout,
                                int64_t ldq, void* Rout, int64_t ldr, int flags) {
  const int k = min(m, n);
  const size_t nq = (size_t)m * k, nr = (size_t)k * n;
  const bool qadj = (flags & TENET_QR_Q_ADJ) != 0, radj = (flags & TENET_QR_R_ADJ) != 0;
  S2* Qo = static_cast<S2*>(Qout);
  S2* Ro = static_cast<S2*>(Rout);
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nq + nr; e += (size_t)gridDim.x * blockDim.x) {
    if (e < nq) {
      const int r = (int)(e % m), c = (int)(e / m);
      const double2 v = reinterpret_cast<const double2*>(Q)[e];
      S2 o;
      o.x = v.x;
      o.y = qadj ? -v.y : v.y;
      if (qadj)
        Qo[(size_t)r * ldq + c] = o;
      else
        Qo[(size_t)c * ldq + r] = o;
    } else {
      const size_t f = e - nq;
      const int r = (int)(f % k), c = (int)(f / k);
      double2 v;
      v.x = 0.0;
      v.y = 0.0;
      if (r <= c) v = reinterpret_cast<const double2*>(W)[(size_t)c * m + r];
      S2 o;
      o.x = v.x;
      o.y = radj ? -v.y : v.y;
      if (radj)
        Ro[(size_t)r * ldr + c] = o;
      else
        Ro[(size_t)c * ldr + r] = o;
    }
  }
}

bool qr_blocked_fits(int64_t m) { return m <= QP_MAX_ROWS; }

// Row-chunked application of one block reflector, M <- (I - V T' V^H) M on
// rows [j0, m) and columns [col_begin, col_end): two launches, each with a
// workgroup per (16-column block, 64-row chunk), so a 512-row update runs on
// 8x the workgroups of the one-pass kernel above.
//   phase Y: Ypart[chunk][block] = V[chunk]^H M[chunk, block]   (16 x 16)
//   phase X: Z = T' sum_chunks Ypart (each workgroup, fixed order);
//            M[chunk, block] -= V[chunk] Z
constexpr int QA_RC = 64;

__device__ __forceinline__ void qa_load_v(const double2* __restrict__ Wc, int m, int j0, int jb, int rb, int rows,
                                          double (&vr)[QA_RC][QB + 1], double (&vi)[QA_RC][QB + 1]) {
  for (int e = threadIdx.x; e < QA_RC * QB; e += QR_THREADS) {
    const int r = e % QA_RC, cc = e / QA_RC;
    const int gr = rb + r, gc = j0 + cc;
    double v_r = 0.0, v_i = 0.0;
    if (r < rows && cc < jb) {
      if (gr == gc) {
        v_r = 1.0;
      } else if (gr > gc) {
        const double2 w = Wc[(size_t)gc * m + gr];
        v_r = w.x;
        v_i = w.y;
      }
    }
    vr[r][cc] = v_r;
    vi[r][cc] = v_i;
  }
}

__device__ __forceinline__ void qa_y_body(const double* __restrict__ W, const double* __restrict__ M,
                                          double* __restrict__ Yp, int m, int ldm, int j0, int jb, int col_begin,
                                          int col_end, int blk, int chunk, int nblk) {
  __shared__ double vr[QA_RC][QB + 1], vi[QA_RC][QB + 1];
  __shared__ double xr[QA_RC][QB + 1], xi[QA_RC][QB + 1];
  __shared__ double yred[4][QB * QB][2];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int cb0 = col_begin + blk * QB;
  const int nc = min(QB, col_end - cb0);
  const int rb = j0 + chunk * QA_RC;
  const int rows = min(QA_RC, m - rb);
  if (nc <= 0 || rows <= 0) return;
  const double2* Wc = reinterpret_cast<const double2*>(W);
  const double2* Mc = reinterpret_cast<const double2*>(M);
  qa_load_v(Wc, m, j0, jb, rb, rows, vr, vi);
  for (int e = tid; e < QA_RC * QB; e += QR_THREADS) {
    const int r = e % QA_RC, cc = e / QA_RC;
    double x_r = 0.0, x_i = 0.0;
    if (r < rows && cc < nc) {
      const double2 x = Mc[(size_t)(cb0 + cc) * ldm + rb + r];
      x_r = x.x;
      x_i = x.y;
    }
    xr[r][cc] = x_r;
    xi[r][cc] = x_i;
  }
  __syncthreads();
  // thread: row group g (rows g, g+16, g+32, g+48) and a 4 x 4 block of Y
  const int g = tid / 16, a0 = 4 * ((tid % 16) / 4), c0 = 4 * (tid % 4);
  double yr_[4][4], yi_[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      yr_[i][j] = 0.0;
      yi_[i][j] = 0.0;
    }
#pragma unroll
  for (int rr = 0; rr < QA_RC / 16; ++rr) {
    const int r = g + 16 * rr;
    double pr[4], pi[4], qr[4], qi[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      pr[i] = vr[r][a0 + i];
      pi[i] = vi[r][a0 + i];
      qr[i] = xr[r][c0 + i];
      qi[i] = xi[r][c0 + i];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {  // conj(V[r][a]) X[r][c]
        yr_[i][j] += pr[i] * qr[j] + pi[i] * qi[j];
        yi_[i][j] += pr[i] * qi[j] - pi[i] * qr[j];
      }
  }
  // row groups of one wave: lanes l, l^16, l^32, l^48
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      yr_[i][j] = xsum32_d(xsum16_d(yr_[i][j]));
      yi_[i][j] = xsum32_d(xsum16_d(yi_[i][j]));
    }
  if (lane < 16) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        yred[wave][(a0 + i) * QB + c0 + j][0] = yr_[i][j];
        yred[wave][(a0 + i) * QB + c0 + j][1] = yi_[i][j];
      }
  }
  __syncthreads();
  // Ypart[chunk][blk][a][c], fixed summation order over the four waves
  double* out = Yp + ((size_t)chunk * nblk + blk) * QB * QB * 2;
  {
    const int e = tid;  // QB * QB = 256 = QR_THREADS
    out[2 * e] = yred[0][e][0] + yred[1][e][0] + yred[2][e][0] + yred[3][e][0];
    out[2 * e + 1] = yred[0][e][1] + yred[1][e][1] + yred[2][e][1] + yred[3][e][1];
  }
}

__global__ __launch_bounds__(QR_THREADS) void qr_apply_y_kernel(const double* __restrict__ W,
                                                                 const double* __restrict__ M, double* __restrict__ Yp,
                                                                 int m, int ldm, int j0, int jb, int col_begin,
                                                                 int col_end) {
  qa_y_body(W, M, Yp, m, ldm, j0, jb, col_begin, col_end, blockIdx.x, blockIdx.y, gridDim.x);
}

__device__ __forceinline__ void qa_x_body(const double* __restrict__ W, const double* __restrict__ T,
                                          const double* __restrict__ Yp, double* __restrict__ M, int m, int ldm,
                                          int j0, int jb, int col_begin, int col_end, int nchunks, int adj, int blk,
                                          int chunk, int nblk) {
  __shared__ double vr[QA_RC][QB + 1], vi[QA_RC][QB + 1];
  __shared__ double yr[QB][QB + 1], yi[QB][QB + 1];
  __shared__ double zr[QB][QB + 1], zi[QB][QB + 1];
  const int tid = threadIdx.x;
  const int cb0 = col_begin + blk * QB;
  const int nc = min(QB, col_end - cb0);
  const int rb = j0 + chunk * QA_RC;
  const int rows = min(QA_RC, m - rb);
  if (nc <= 0 || rows <= 0) return;
  const double2* Wc = reinterpret_cast<const double2*>(W);
  double2* Mc = reinterpret_cast<double2*>(M);
  qa_load_v(Wc, m, j0, jb, rb, rows, vr, vi);
  {
    // Y = sum over chunks (fixed order), one entry per thread
    const int a = tid / QB, c = tid % QB;
    double sr = 0.0, si = 0.0;
    for (int ch = 0; ch < nchunks; ++ch) {
      const double* in = Yp + ((size_t)ch * nblk + blk) * QB * QB * 2;
      sr += in[2 * (a * QB + c)];
      si += in[2 * (a * QB + c) + 1];
    }
    yr[a][c] = sr;
    yi[a][c] = si;
  }
  __syncthreads();
  {
    // Z = T' Y
    const int a = tid / QB, c = tid % QB;
    double sr = 0.0, si = 0.0;
    for (int b = 0; b < jb; ++b) {
      double tr, ti;
      if (adj) {
        tr = T[2 * (a * QB + b)];
        ti = -T[2 * (a * QB + b) + 1];
      } else {
        tr = T[2 * (b * QB + a)];
        ti = T[2 * (b * QB + a) + 1];
      }
      sr += tr * yr[b][c] - ti * yi[b][c];
      si += tr * yi[b][c] + ti * yr[b][c];
    }
    zr[a][c] = (a < jb) ? sr : 0.0;
    zi[a][c] = (a < jb) ? si : 0.0;
  }
  __syncthreads();
  // M[chunk rows] -= V Z: thread owns row tid % 64 and 4 columns
  const int rsub = tid % QA_RC, cq = 4 * (tid / QA_RC);
  if (rsub < rows) {
    double sr[4] = {0, 0, 0, 0}, si[4] = {0, 0, 0, 0};
#pragma unroll
    for (int b = 0; b < QB; ++b) {
      const double pr = vr[rsub][b], pi = vi[rsub][b];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        sr[j] += pr * zr[b][cq + j] - pi * zi[b][cq + j];
        si[j] += pr * zi[b][cq + j] + pi * zr[b][cq + j];
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (cq + j < nc) {
        double2 x = Mc[(size_t)(cb0 + cq + j) * ldm + rb + rsub];
        x.x -= sr[j];
        x.y -= si[j];
        Mc[(size_t)(cb0 + cq + j) * ldm + rb + rsub] = x;
      }
    }
  }
}

__global__ __launch_bounds__(QR_THREADS) void qr_apply_x_kernel(const double* __restrict__ W,
                                                                 const double* __restrict__ T,
                                                                 const double* __restrict__ Yp, double* __restrict__ M,
                                                                 int m, int ldm, int j0, int jb, int col_begin,
                                                                 int col_end, int nchunks, int adj) {
  qa_x_body(W, T, Yp, M, m, ldm, j0, jb, col_begin, col_end, nchunks, adj, blockIdx.x, blockIdx.y, gridDim.x);
}

// Batched form: Z = T^H sum_chunks Ypart formed once per (problem, column
// block) and stored over that block's chunk-0 partial (its only reader is
// the workgroup that overwrites it, after the sum), so each row-chunk
// workgroup of the update reads one 16 x 16 block instead of every partial.
__device__ __forceinline__ void qa_z_body(const double* __restrict__ T, double* __restrict__ Yp, int jb,
                                          int nchunks, int blk, int nblk) {
  __shared__ double yr[QB][QB + 1], yi[QB][QB + 1];
  const int tid = threadIdx.x;
  const int a = tid / QB, c = tid % QB;
  {
    double sr = 0.0, si = 0.0;
    for (int ch = 0; ch < nchunks; ++ch) {
      const double* in = Yp + ((size_t)ch * nblk + blk) * QB * QB * 2;
      sr += in[2 * (a * QB + c)];
      si += in[2 * (a * QB + c) + 1];
    }
    yr[a][c] = sr;
    yi[a][c] = si;
  }
  __syncthreads();
  double sr = 0.0, si = 0.0;
  for (int b = 0; b < jb; ++b) {
    const double tr = T[2 * (a * QB + b)], ti = -T[2 * (a * QB + b) + 1];
    sr += tr * yr[b][c] - ti * yi[b][c];
    si += tr * yi[b][c] + ti * yr[b][c];
  }
  double* out = Yp + (size_t)blk * QB * QB * 2;
  out[2 * (a * QB + c)] = (a < jb) ? sr : 0.0;
  out[2 * (a * QB + c) + 1] = (a < jb) ? si : 0.0;
}

__device__ __forceinline__ void qa_xz_body(const double* __restrict__ W, const double* __restrict__ Z,
                                           double* __restrict__ M, int m, int ldm, int j0, int jb, int col_begin,
                                           int col_end, int blk, int chunk) {
  __shared__ double vr[QA_RC][QB + 1], vi[QA_RC][QB + 1];
  __shared__ double zr[QB][QB + 1], zi[QB][QB + 1];
  const int tid = threadIdx.x;
  const int cb0 = col_begin + blk * QB;
  const int nc = min(QB, col_end - cb0);
  const int rb = j0 + chunk * QA_RC;
  const int rows = min(QA_RC, m - rb);
  if (nc <= 0 || rows <= 0) return;
  const double2* Wc = reinterpret_cast<const double2*>(W);
  double2* Mc = reinterpret_cast<double2*>(M);
  qa_load_v(Wc, m, j0, jb, rb, rows, vr, vi);
  {
    const int a = tid / QB, c = tid % QB;
    zr[a][c] = Z[2 * (a * QB + c)];
    zi[a][c] = Z[2 * (a * QB + c) + 1];
  }
  __syncthreads();
  // M[chunk rows] -= V Z: thread owns row tid % 64 and 4 columns
  const int rsub = tid % QA_RC, cq = 4 * (tid / QA_RC);
  if (rsub < rows) {
    double sr[4] = {0, 0, 0, 0}, si[4] = {0, 0, 0, 0};
#pragma unroll
    for (int b = 0; b < QB; ++b) {
      const double pr = vr[rsub][b], pi = vi[rsub][b];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        sr[j] += pr * zr[b][cq + j] - pi * zi[b][cq + j];
        si[j] += pr * zi[b][cq + j] + pi * zr[b][cq + j];
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (cq + j < nc) {
        double2 x = Mc[(size_t)(cb0 + cq + j) * ldm + rb + rsub];
        x.x -= sr[j];
        x.y -= si[j];
        Mc[(size_t)(cb0 + cq + j) * ldm + rb + rsub] = x;
      }
    }
  }
}

// Batched trailing update after panel p: blockIdx.z is the problem; the
// problem's trailing columns [j0 + jb, n) of W itself, rows [j0, m).  Ypart
// strides by the grid's (largest) column-block count.
// (columns end at min(n, cap): the panels of a 64-column block update only
// the rest of their block; the block reflector then updates the columns
// right of the block at once)
__global__ __launch_bounds__(QR_THREADS) void qr_apply_y_batched_kernel(const QrBatchProb* __restrict__ probs, int p,
                                                                         int cap) {
  const QrBatchProb& q = probs[blockIdx.z];
  const int k = min(q.m, q.n), j0 = p * QB;
  if (j0 >= k) return;
  const int jb = min(QB, k - j0);
  qa_y_body(q.W, q.W, q.Yp, q.m, q.m, j0, jb, j0 + jb, min(q.n, cap), blockIdx.x, blockIdx.y, gridDim.x);
}

__global__ __launch_bounds__(QR_THREADS) void qr_apply_z_batched_kernel(const QrBatchProb* __restrict__ probs, int p,
                                                                         int cap) {
  const QrBatchProb& q = probs[blockIdx.z];
  const int k = min(q.m, q.n), j0 = p * QB;
  if (j0 >= k) return;
  const int jb = min(QB, k - j0);
  if (j0 + jb + (int)blockIdx.x * QB >= min(q.n, cap)) return;  // no such column block
  const int nch = (q.m - j0 + QA_RC - 1) / QA_RC;
  qa_z_body(q.T + (size_t)p * QB * QB * 2, q.Yp, jb, nch, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(QR_THREADS) void qr_apply_x_batched_kernel(const QrBatchProb* __restrict__ probs, int p,
                                                                         int cap) {
  const QrBatchProb& q = probs[blockIdx.z];
  const int k = min(q.m, q.n), j0 = p * QB;
  if (j0 >= k) return;
  const int jb = min(QB, k - j0);
  qa_xz_body(q.W, q.Yp + (size_t)blockIdx.x * QB * QB * 2, q.W, q.m, q.m, j0, jb, j0 + jb, min(q.n, cap),
             blockIdx.x, blockIdx.y);
}

// ---- block reflectors of 64 columns (four panels) ----------------------------
// After the four panels of block [b0, b0 + kb) are factored (each updating
// only the rest of the block), the columns right of the block get the whole
// block reflector H_b0 ... H_{b0+kb-1} = I - V T V^H at once: one pass over
// the trailing matrix per 64 columns instead of per 16.
constexpr int QBW = 64;  // block width
constexpr int WRC = 64;  // rows per chunk of the block update
constexpr int WRH = 32;  // rows per LDS pass

// V(r, l) of the block starting at column b0: 0 above row b0 + l, 1 on it,
// the stored reflector below
__device__ __forceinline__ double2 blk_v(const double2* __restrict__ W, int m, int b0, int r, int l) {
  const int c = b0 + l;
  if (r < c) return make_double2(0.0, 0.0);
  if (r == c) return make_double2(1.0, 0.0);
  return W[(size_t)c * m + r];
}

// Cross products of the block's reflectors, S = V^H V for panel(l) < panel(j),
// summed per 64-row chunk (blockIdx.x) into partials Sp[chunk] (64 x 64,
// the trailing update's partial buffer, free at this point); each thread
// sums a 4 x 4 block.
__global__ __launch_bounds__(QR_THREADS) void qr_sblock_batched_kernel(const QrBatchProb* __restrict__ probs, int b0) {
  __shared__ double2 vs[WRH][QBW + 1];
  const QrBatchProb& q = probs[blockIdx.y];
  const int k = min(q.m, q.n);
  if (b0 >= k) return;
  const int kb = min(QBW, k - b0);
  if (b0 + kb >= q.n) return;  // no columns right of the block
  const int rb = b0 + (int)blockIdx.x * WRC;
  if (rb >= q.m) return;
  const double2* W = reinterpret_cast<const double2*>(q.W);
  const int tid = threadIdx.x;
  const int lb = tid / 16, jb = tid % 16;
  const bool mine = (lb / 4) < (jb / 4) && 4 * jb < kb;
  double sr[4][4], si[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      sr[a][c] = 0.0;
      si[a][c] = 0.0;
    }
  for (int h = 0; h < WRC; h += WRH) {
    const int r0 = rb + h;
    __syncthreads();
    for (int e = tid; e < WRH * QBW; e += QR_THREADS) {
      const int rr = e % WRH, l = e / WRH, r = r0 + rr;
      vs[rr][l] = (r < q.m && l < kb) ? blk_v(W, q.m, b0, r, l) : make_double2(0.0, 0.0);
    }
    __syncthreads();
    if (mine) {
      for (int rr = 0; rr < WRH; ++rr) {
        double2 va[4], vb[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          va[a] = vs[rr][4 * lb + a];
          vb[a] = vs[rr][4 * jb + a];
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int c = 0; c < 4; ++c) {  // conj(v_l) v_j
            sr[a][c] += va[a].x * vb[c].x + va[a].y * vb[c].y;
            si[a][c] += va[a].x * vb[c].y - va[a].y * vb[c].x;
          }
      }
    }
  }
  if (mine) {
    double* out = q.Yp + (size_t)blockIdx.x * QBW * QBW * 2;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int f = (4 * lb + a) * QBW + 4 * jb + c;
        out[2 * f] = sr[a][c];
        out[2 * f + 1] = si[a][c];
      }
  }
}

// T of the block, assembled from the panels' own 16 x 16 factors: for
// H_A H_B with V = [V_A V_B], T = [[T_A, -T_A (V_A^H V_B) T_B], [0, T_B]]
// (the rule zlarft's recursion satisfies), applied for B = panel 1, 2, 3 of
// the block in turn, with V_A^H V_B the chunk partials of
// qr_sblock_batched_kernel summed in chunk order; one workgroup per problem.
// TB (column-major, 64 x 64) <- T.
__global__ __launch_bounds__(QR_THREADS) void qr_tblock_batched_kernel(const QrBatchProb* __restrict__ probs, int b0) {
  __shared__ double2 S[QBW][QBW + 1];  // S[l][j] = v_l^H v_j, panel(l) < panel(j)
  __shared__ double2 T[QBW][QBW + 1];  // T[i][j]
  __shared__ double2 Ys[QBW - QB][QB + 1];
  const QrBatchProb& q = probs[blockIdx.x];
  const int k = min(q.m, q.n);
  if (b0 >= k) return;
  const int kb = min(QBW, k - b0);
  if (b0 + kb >= q.n) return;  // no columns right of the block
  const int tid = threadIdx.x;
  const int nch = (q.m - b0 + WRC - 1) / WRC;
  const int p0 = b0 / QB;
  for (int e = tid; e < QBW * QBW; e += QR_THREADS) {
    const int l = e / QBW, j = e % QBW;
    double2 sv = make_double2(0.0, 0.0);
    if (l / QB < j / QB && j < kb)
      for (int ch = 0; ch < nch; ++ch) {
        const double* in = q.Yp + (size_t)ch * QBW * QBW * 2;
        sv.x += in[2 * e];
        sv.y += in[2 * e + 1];
      }
    S[l][j] = sv;
    // the diagonal blocks: each panel's T (column-major 16 x 16), zero elsewhere
    const int i2 = e % QBW, j2 = e / QBW;
    double2 v = make_double2(0.0, 0.0);
    if (i2 / QB == j2 / QB && j2 < kb) {
      const double* Tp = q.T + (size_t)(p0 + j2 / QB) * QB * QB * 2;
      const int f = (j2 % QB) * QB + (i2 % QB);
      v = make_double2(Tp[2 * f], Tp[2 * f + 1]);
    }
    T[i2][j2] = v;
  }
  __syncthreads();
  for (int bq = 1; bq * QB < kb; ++bq) {
    const int na = bq * QB, nbw = min(QB, kb - na);  // A = [0, na), B = [na, na + nbw)
    // Y = T_A S(A, B)
    for (int e = tid; e < na * QB; e += QR_THREADS) {
      const int i = e % na, c = e / na;
      double yr = 0.0, yi = 0.0;
      if (c < nbw)
        for (int l = i; l < na; ++l) {
          const double2 t = T[i][l], x = S[l][na + c];
          yr += t.x * x.x - t.y * x.y;
          yi += t.x * x.y + t.y * x.x;
        }
      Ys[i][c] = make_double2(yr, yi);
    }
    __syncthreads();
    // T(A, B) = -Y T_B
    for (int e = tid; e < na * QB; e += QR_THREADS) {
      const int i = e % na, c = e / na;
      if (c < nbw) {
        double tr = 0.0, ti = 0.0;
        for (int l = 0; l <= c; ++l) {  // T_B upper triangular
          const double2 y = Ys[i][l], t = T[na + l][na + c];
          tr += y.x * t.x - y.y * t.y;
          ti += y.x * t.y + y.y * t.x;
        }
        T[i][na + c] = make_double2(-tr, -ti);
      }
    }
    __syncthreads();
  }
  for (int e = tid; e < QBW * QBW; e += QR_THREADS) {
    const int i = e % QBW, j = e / QBW;
    const double2 v = (i < kb && j < kb) ? T[i][j] : make_double2(0.0, 0.0);
    q.TB[2 * e] = v.x;
    q.TB[2 * e + 1] = v.y;
  }
}

// The block reflector on the columns right of the block in one launch, on
// FP64 MFMA (v_mfma_f64_16x16x4): per workgroup 32 columns of C (the trailing
// columns of W, rows b0 .. m) and
//   Y = V^H C (64 x 32),  Z = T^H Y,  C <- C - V Z,
// i.e. C <- (I - V T^H V^H) C.  Y: the rows in 32-row passes staged through
// LDS (V transposed for the A operand); Z: T^H read from TB (L2) as the A
// operand, Y from LDS; C - V Z: V read as the A operand straight from W,
// Z from LDS, C read and written from the accumulators.  Every complex
// product is three real ones (Gauss); wave w owns output rows 16 w .. 16 w + 15
// of each 64-row tile and both 16-column halves.
typedef double qr_d4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ qr_d4v qr_mfma(double a, double b, qr_d4v c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
constexpr int WF_COLS = 32;  // trailing columns per workgroup
constexpr int WF_RP = 32;    // rows per staged pass of Y
struct WfSmem {
  union {
    struct {
      double vr[WF_RP][QBW + 1], vi[WF_RP][QBW + 1];
      double cr[WF_RP][WF_COLS + 1], ci[WF_RP][WF_COLS + 1];
    } y;
    struct {
      double zr[QBW][WF_COLS + 1], zi[QBW][WF_COLS + 1];
    } z;
  };
};

__global__ __launch_bounds__(QR_THREADS, 2) void qr_wfused_batched_kernel(const QrBatchProb* __restrict__ probs, int b0) {
  __shared__ WfSmem sm;
  const QrBatchProb& q = probs[blockIdx.y];
  const int k = min(q.m, q.n);
  if (b0 >= k) return;
  const int kb = min(QBW, k - b0);
  const int cb0 = b0 + kb + (int)blockIdx.x * WF_COLS;
  if (cb0 >= q.n) return;
  const int nc = min(WF_COLS, q.n - cb0);
  const int m = q.m;
  double2* W = reinterpret_cast<double2*>(q.W);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  // ---- Y = V^H C: T1 = Vr^T Cr, T2 = Vi^T Ci, T3 = (Vr + Vi)^T (Cr - Ci);
  //      Re Y = T1 + T2, Im Y = T1 - T2 - T3
  qr_d4v t1[2], t2[2], t3[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) t1[h] = t2[h] = t3[h] = qr_d4v{0, 0, 0, 0};
  // the next pass's V and C entries are loaded into registers while the
  // current pass multiplies (one memory latency per pass, not one per load)
  constexpr int NV = WF_RP * QBW / QR_THREADS, NCC = WF_RP * WF_COLS / QR_THREADS;
  double2 pv[NV], pc[NCC];
  auto load_pass = [&](int r0) {
#pragma unroll
    for (int u = 0; u < NV; ++u) {
      const int e = tid + u * QR_THREADS;
      const int rr = e % WF_RP, l = e / WF_RP, r = r0 + rr;
      pv[u] = (r < m && l < kb) ? blk_v(W, m, b0, r, l) : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < NCC; ++u) {
      const int e = tid + u * QR_THREADS;
      const int rr = e % WF_RP, c = e / WF_RP, r = r0 + rr;
      pc[u] = (r < m && c < nc) ? W[(size_t)(cb0 + c) * m + r] : make_double2(0.0, 0.0);
    }
  };
  load_pass(b0);
  for (int r0 = b0; r0 < m; r0 += WF_RP) {
    __syncthreads();
#pragma unroll
    for (int u = 0; u < NV; ++u) {
      const int e = tid + u * QR_THREADS;
      sm.y.vr[e % WF_RP][e / WF_RP] = pv[u].x;
      sm.y.vi[e % WF_RP][e / WF_RP] = pv[u].y;
    }
#pragma unroll
    for (int u = 0; u < NCC; ++u) {
      const int e = tid + u * QR_THREADS;
      sm.y.cr[e % WF_RP][e / WF_RP] = pc[u].x;
      sm.y.ci[e % WF_RP][e / WF_RP] = pc[u].y;
    }
    __syncthreads();
    if (r0 + WF_RP < m) load_pass(r0 + WF_RP);
#pragma unroll 2
    for (int st = 0; st < WF_RP / 4; ++st) {
      const int row = 4 * st + lk;
      const double ar = sm.y.vr[row][16 * w + li], ai = sm.y.vi[row][16 * w + li];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const double br = sm.y.cr[row][16 * h + li], bi = sm.y.ci[row][16 * h + li];
        t1[h] = qr_mfma(ar, br, t1[h]);
        t2[h] = qr_mfma(ai, bi, t2[h]);
        t3[h] = qr_mfma(ar + ai, br - bi, t3[h]);
      }
    }
  }
  // Y into LDS: row 16 w + 4 j + lk, column 16 h + li
  __syncthreads();
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      sm.z.zr[16 * w + 4 * j + lk][16 * h + li] = t1[h][j] + t2[h][j];
      sm.z.zi[16 * w + 4 * j + lk][16 * h + li] = t1[h][j] - t2[h][j] - t3[h][j];
    }
  __syncthreads();
  // ---- Z = T^H Y: A = conj(T)^T (T upper triangular, column-major in TB),
  //      Re Z = T1 + T2, Im Z = T1 - T2 - T3 as above
#pragma unroll
  for (int h = 0; h < 2; ++h) t1[h] = t2[h] = t3[h] = qr_d4v{0, 0, 0, 0};
  const int a = 16 * w + li;  // Z row of this lane's A operand
  const double2* TB2 = reinterpret_cast<const double2*>(q.TB);
  for (int s0 = 0; s0 < QBW / 4; s0 += 8) {  // eight K steps' T entries in flight
    double2 tv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int bb = 4 * (s0 + u) + lk;
      tv[u] = (bb <= a && a < kb) ? TB2[bb + QBW * a] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int bb = 4 * (s0 + u) + lk;
      const double tr = tv[u].x, ti = tv[u].y;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const double yr = sm.z.zr[bb][16 * h + li], yi = sm.z.zi[bb][16 * h + li];
        t1[h] = qr_mfma(tr, yr, t1[h]);
        t2[h] = qr_mfma(ti, yi, t2[h]);
        t3[h] = qr_mfma(tr + ti, yr - yi, t3[h]);
      }
    }
  }
  __syncthreads();  // every wave has read Y
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      sm.z.zr[16 * w + 4 * j + lk][16 * h + li] = t1[h][j] + t2[h][j];
      sm.z.zi[16 * w + 4 * j + lk][16 * h + li] = t1[h][j] - t2[h][j] - t3[h][j];
    }
  __syncthreads();
  // ---- C <- C - V Z over 64-row tiles: T1 = Vr Zr, T2 = Vi Zi,
  //      T3 = (Vr + Vi)(Zr + Zi); Re = T1 - T2, Im = T3 - T1 - T2
  for (int r0 = b0; r0 < m; r0 += QBW) {
    const int r = r0 + 16 * w + li;  // this lane's V row (A operand)
#pragma unroll
    for (int h = 0; h < 2; ++h) t1[h] = t2[h] = t3[h] = qr_d4v{0, 0, 0, 0};
    for (int s0 = 0; s0 < QBW / 4; s0 += 8) {  // eight K steps' V entries in flight
      double2 vv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int l = 4 * (s0 + u) + lk;
        vv[u] = (r < m && l < kb) ? blk_v(W, m, b0, r, l) : make_double2(0.0, 0.0);
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int l = 4 * (s0 + u) + lk;
        const double2 v = vv[u];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const double zr = sm.z.zr[l][16 * h + li], zi = sm.z.zi[l][16 * h + li];
          t1[h] = qr_mfma(v.x, zr, t1[h]);
          t2[h] = qr_mfma(v.y, zi, t2[h]);
          t3[h] = qr_mfma(v.x + v.y, zr + zi, t3[h]);
        }
      }
    }
    // accumulator j of this lane: row r0 + 16 w + 4 j + lk, column 16 h + li
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int c = 16 * h + li;
      if (c >= nc) continue;
      double2* col = W + (size_t)(cb0 + c) * m;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int rr = r0 + 16 * w + 4 * j + lk;
        if (rr < m) {
          double2 x = col[rr];
          x.x -= t1[h][j] - t2[h][j];
          x.y -= t3[h][j] - t1[h][j] - t2[h][j];
          col[rr] = x;
        }
      }
    }
  }
}

// M <- (I - V T' V^H) M (adj = 1: T' = T^H, forward application during the
// factorisation; adj = 0: T' = T, backward accumulation of Q)
static int qr_apply_chunked(const double* W, const double* T, double* M, int m, int ldm, int j0, int jb, int c_begin,
                            int c_end, int adj, double* Yp, hipStream_t s) {
  if (c_end <= c_begin || m <= j0) return TENET_OK;
  const unsigned nblk = (unsigned)((c_end - c_begin + QB - 1) / QB);
  const unsigned nch = (unsigned)((m - j0 + QA_RC - 1) / QA_RC);
  hipLaunchKernelGGL(qr_apply_y_kernel, dim3(nblk, nch), dim3(QR_THREADS), 0, s, W, M, Yp, m, ldm, j0, jb, c_begin,
                     c_end);
  TENET_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(qr_apply_x_kernel, dim3(nblk, nch), dim3(QR_THREADS), 0, s, W, T, Yp, M, m, ldm, j0, jb, c_begin,
                     c_end, (int)nch, adj);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

// Householder QR of every problem of a batch in place (R on and above the
// diagonal of W, the reflectors below; tau and the panels' T factors kept):
// per panel one launch of the panel kernel over the problems and one pair of
// trailing-update launches.  Every m <= 512 (the register panel).
size_t qr_batched_yp_bytes(int64_t m, int64_t max_n) {
  // the reflector block's cross products (64 x 64 per 64-row chunk) and the
  // panels' trailing-update partials (16 x 16 per chunk and column block)
  return (size_t)((m + WRC - 1) / WRC) * QBW * QBW * 16 +
         (size_t)((m + QA_RC - 1) / QA_RC) * (size_t)((max_n + QB - 1) / QB) * QB * QB * 16;
}
size_t qr_batched_tb_bytes() { return (size_t)QBW * QBW * 16; }
size_t qr_batched_t_bytes(int64_t m, int64_t n) { return (size_t)((std::min(m, n) + QB - 1) / QB) * QB * QB * 16; }

int launch_qr_batched(const QrBatchProb* probs, int64_t nprob, int max_m, int max_n, int max_k, hipStream_t s) {
  if (nprob <= 0 || max_k <= 0) return TENET_OK;
  if (max_m > 512 || nprob > 65535) {
    set_error("batched QR: rows above 512 or more than 65535 problems");
    return TENET_ERR_UNSUPPORTED;
  }
  for (int b0 = 0; b0 < max_k; b0 += QBW) {
    const int cap = b0 + QBW;
    for (int j0 = b0; j0 < std::min(cap, max_k); j0 += QB) {
      const int p = j0 / QB;
      const int mr = max_m - j0;
      const dim3 g1((unsigned)nprob);
      if (mr <= 64)
        hipLaunchKernelGGL(qr_panel_v4_batched_kernel<1>, g1, dim3(64 * QB), 0, s, probs, p);
      else if (mr <= 128)
        hipLaunchKernelGGL(qr_panel_v4_batched_kernel<2>, g1, dim3(64 * QB), 0, s, probs, p);
      else if (mr <= 256)
        hipLaunchKernelGGL(qr_panel_v4_batched_kernel<4>, g1, dim3(64 * QB), 0, s, probs, p);
      else
        hipLaunchKernelGGL(qr_panel_v4_batched_kernel<8>, g1, dim3(64 * QB), 0, s, probs, p);
      TENET_HIP_CHECK(hipGetLastError());
      // the rest of the block: a problem's columns start at j0 + jb >= j0 + 1,
      // so blocks over [j0 + 1, min(max_n, cap)) cover every problem's
      const int ce = std::min(max_n, cap);
      if (ce - j0 <= 1) continue;
      const dim3 g2((unsigned)((ce - j0 - 1 + QB - 1) / QB), (unsigned)((mr + QA_RC - 1) / QA_RC), (unsigned)nprob);
      hipLaunchKernelGGL(qr_apply_y_batched_kernel, g2, dim3(QR_THREADS), 0, s, probs, p, cap);
      TENET_HIP_CHECK(hipGetLastError());
      hipLaunchKernelGGL(qr_apply_z_batched_kernel, dim3(g2.x, 1, g2.z), dim3(QR_THREADS), 0, s, probs, p, cap);
      TENET_HIP_CHECK(hipGetLastError());
      hipLaunchKernelGGL(qr_apply_x_batched_kernel, g2, dim3(QR_THREADS), 0, s, probs, p, cap);
      TENET_HIP_CHECK(hipGetLastError());
    }
    // the block reflector on the columns right of the block (a problem's
    // start at b0 + kb >= b0 + 1)
    if (max_n - b0 <= 1) continue;
    const unsigned nchunk = (unsigned)((max_m - b0 + WRC - 1) / WRC);
    hipLaunchKernelGGL(qr_sblock_batched_kernel, dim3(nchunk, (unsigned)nprob), dim3(QR_THREADS), 0, s, probs, b0);
    TENET_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(qr_tblock_batched_kernel, dim3((unsigned)nprob), dim3(QR_THREADS), 0, s, probs, b0);
    TENET_HIP_CHECK(hipGetLastError());
    const dim3 gf((unsigned)((max_n - b0 - 1 + WF_COLS - 1) / WF_COLS), (unsigned)nprob);
    hipLaunchKernelGGL(qr_wfused_batched_kernel, gf, dim3(QR_THREADS), 0, s, probs, b0);
    TENET_HIP_CHECK(hipGetLastError());
  }
  return TENET_OK;
}

size_t qr_blocked_ws_bytes(int64_t m, int64_t n) {
  const int64_t k = std::min(m, n);
  const int64_t np = (k + QB - 1) / QB;
  const int64_t ych = (m + QA_RC - 1) / QA_RC, yblk = (std::max(n, k) + QB - 1) / QB;
  return (size_t)m * n * 16 + (size_t)m * k * 16 + (size_t)k * 16 + (size_t)np * QB * QB * 16 +
         (size_t)ych * yblk * QB * QB * 16 + 2048;
}

// ---------------------------------------------------------------------------
// Register-resident QR for m, n <= 64 (the canonize! moves at small bond
// dimensions: 64 x 32, 2 x 32, ...): one workgroup, A read in place (adjoint
// and complex64 input converted on the load), no workspace.  With RL rows per
// lane (RL = 4 or 8), lane l of wave w holds rows RL rb .. RL rb + RL - 1,
// rb = l / RL, of column RL w + (l % RL) in registers: a wave's RL columns
// are updated by the same instructions and a column's v^H a is an in-lane
// sum over RL rows plus the exchange levels over the 64 / RL row blocks.
// Step j: every wave applies reflector j (v, tau from LDS) to its columns
// right of j; the wave owning column j + 1 then forms reflector j + 1 (every
// lane of it computes the scalars; no lane-serial section) and publishes it:
// one barrier per reflector.  Q = H_1 ... H_k [I; 0] is accumulated
// afterwards without barriers (all reflectors are in LDS), each wave on its
// own columns.  The reflectors are qr_factor_kernel's (v(0) = 1,
// beta = -sign(Re alpha) ||(alpha, x)||, H^H applied from the left).
// RL = 4 halves each wave's work per step and doubles the waves; it is used
// while the workgroup stays within 1024 threads.
// ---------------------------------------------------------------------------
constexpr int QW_MAX = 64;  // m, n <= 64
constexpr int QW_PROD = 4;  // waves that absorb R into the neighbour

// Row i of a reflector sits at (i % RL) * (64 / RL) + i / RL: the row blocks a
// wave reads with one instruction are then adjacent in LDS.

// Sum over the row blocks of a column: lanes l, l + RL, l + 2 RL, ... of the
// 16-lane row by rotations (row_ror:8, and row_ror:4 first when RL = 4),
// then across rows and halves.
template <int RL>
__device__ __forceinline__ double qw_rowsum(double v) {
  if (RL == 4) v += dpp_d<0x124>(v);  // row_ror:4
  v += dpp_d<0x128>(v);                // row_ror:8
  return xsum32_d(xsum16_d(v));
}

// v^H a over this lane's RL rows (two accumulator chains), then the column sum
template <int RL>
__device__ __forceinline__ void qw_dot(const double (&vr)[RL], const double (&vi)[RL], const double (&xr)[RL],
                                       const double (&xi)[RL], double& dr, double& di) {
  constexpr int H = RL / 2;
  double r0 = 0.0, i0 = 0.0, r1 = 0.0, i1 = 0.0;
#pragma unroll
  for (int r = 0; r < H; ++r) {
    r0 = fma(vr[r], xr[r], fma(vi[r], xi[r], r0));
    i0 = fma(vr[r], xi[r], fma(-vi[r], xr[r], i0));
    r1 = fma(vr[r + H], xr[r + H], fma(vi[r + H], xi[r + H], r1));
    i1 = fma(vr[r + H], xi[r + H], fma(-vi[r + H], xr[r + H], i1));
  }
  dr = qw_rowsum<RL>(r0 + r1);
  di = qw_rowsum<RL>(i0 + i1);
}

// x -= v f over this lane's RL rows
template <int RL>
__device__ __forceinline__ void qw_axpy(const double (&vr)[RL], const double (&vi)[RL], double fr, double fi,
                                        double (&xr)[RL], double (&xi)[RL]) {
#pragma unroll
  for (int r = 0; r < RL; ++r) {
    xr[r] = fma(-vr[r], fr, fma(vi[r], fi, xr[r]));
    xi[r] = fma(-vr[r], fi, fma(-vi[r], fr, xi[r]));
  }
}

template <int RL>
__device__ __forceinline__ void qw_load_v(const double2* Vj, int rb, double (&vr)[RL], double (&vi)[RL]) {
#pragma unroll
  for (int r = 0; r < RL; ++r) {
    const double2 x = Vj[r * (64 / RL) + rb];  // row RL rb + r
    vr[r] = x.x;
    vi[r] = x.y;
  }
}

// Reflector of column jj, called by the whole wave that owns it (wave jj / RL;
// jj uniform); on return that column's registers hold R(:, jj) (beta on the
// diagonal, zeros below) and V / T hold v and tau.  Branch-free apart from
// the stores.
template <int RL>
__device__ __forceinline__ void qw_reflector(double (&ar)[RL], double (&ai)[RL], int jj, int cl, int rb,
                                             double2 (*V)[QW_MAX], double2* T) {
  constexpr int H = RL / 2;
  const int jr = jj % RL;
  double p0 = 0.0, p1 = 0.0;
#pragma unroll
  for (int r = 0; r < H; ++r) {
    const double t0 = fma(ar[r], ar[r], ai[r] * ai[r]);
    const double t1 = fma(ar[r + H], ar[r + H], ai[r + H] * ai[r + H]);
    p0 += (RL * rb + r > jj) ? t0 : 0.0;
    p1 += (RL * rb + r + H > jj) ? t1 : 0.0;
  }
  const double part = qw_rowsum<RL>(p0 + p1);
  double sr = ar[0], si = ai[0];
#pragma unroll
  for (int r = 1; r < RL; ++r) {
    sr = (r == jr) ? ar[r] : sr;
    si = (r == jr) ? ai[r] : si;
  }
  const double xn = readlane_d(part, jr);  // ||x||^2 of column jj (lane jr: row block 0 of that column)
  const double alr = readlane_d(sr, jj), ali = readlane_d(si, jj);  // A(jj, jj) sits in lane jj
  double tr = 0.0, ti = 0.0, scr = 0.0, sci = 0.0, beta = alr;
  if (!(xn == 0.0 && ali == 0.0)) {
    const double nrm = sqrt(fma(alr, alr, fma(ali, ali, xn)));
    beta = (alr >= 0.0) ? -nrm : nrm;
    const double dr = alr - beta, di = ali;
    const double ib = 1.0 / beta, iden = 1.0 / fma(dr, dr, di * di);
    tr = (beta - alr) * ib;
    ti = -ali * ib;
    scr = dr * iden;  // 1 / (alpha - beta)
    sci = -di * iden;
  }
  const bool own = cl == jr;
  double2 v[RL];
#pragma unroll
  for (int r = 0; r < RL; ++r) {
    const int i = RL * rb + r;
    const bool below = i > jj, diag = i == jj;
    v[r].x = below ? fma(ar[r], scr, -ai[r] * sci) : (diag ? 1.0 : 0.0);
    v[r].y = below ? fma(ar[r], sci, ai[r] * scr) : 0.0;
    ar[r] = (own && below) ? 0.0 : ((own && diag) ? beta : ar[r]);
    ai[r] = (own && (below || diag)) ? 0.0 : ai[r];
  }
  if (own) {
#pragma unroll
    for (int r = 0; r < RL; ++r) V[jj][r * (64 / RL) + rb] = v[r];
    if (rb == 0) {
      double2 t;
      t.x = tr;
      t.y = ti;
      T[jj] = t;
    }
  }
}

// Neighbour absorption (generic_canonize_site!'s R = binary_einsum(R, B)):
// side 0: Bo (k x nb) = R B with B n x nb; side 1: Bo (nb x k) = B R^H with
// B nb x n.  B == nullptr: none.
template <typename S2>
struct QwAbsorb {
  const S2* B;
  S2* Bo;
  int64_t ldb, ldbo;
  int nb, side;
};

// Bo from R (Rs[c * 64 + q] = R(q, c), zero below the diagonal), by the
// QW_PROD product waves (thread pt of them) while the column waves form Q.
// A thread owns one neighbour index p and eight consecutive q: each B
// element it loads serves eight outputs.
template <typename S2>
__device__ __forceinline__ void qw_absorb(const QwAbsorb<S2>& ab, const double2* Rs, int n, int k, int pt) {
  const int nb = ab.nb, total = nb * ((k + 7) >> 3);
  for (int e = pt; e < total; e += 64 * QW_PROD) {
    const int p = e % nb, q0 = 8 * (e / nb);
    double sr[8], si[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      sr[t] = 0.0;
      si[t] = 0.0;
    }
    // rows q0 + t > c of R are zero; rows >= k are not stored below
    if (ab.side == 0) {  // Bo(q, p) = sum_c R(q, c) B(c, p)
      const S2* b = ab.B + (size_t)p * ab.ldb;
#pragma unroll 4
      for (int c = q0; c < n; ++c) {
        const S2 x = b[c];
        const double xr = x.x, xi = x.y;
        const double2* r = Rs + c * QW_MAX + q0;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          const double2 y = r[t];
          sr[t] = fma(y.x, xr, fma(-y.y, xi, sr[t]));
          si[t] = fma(y.x, xi, fma(y.y, xr, si[t]));
        }
      }
    } else {  // Bo(p, q) = sum_c B(p, c) conj(R(q, c))
      const S2* b = ab.B + p;
#pragma unroll 4
      for (int c = q0; c < n; ++c) {
        const S2 x = b[(size_t)c * ab.ldb];
        const double xr = x.x, xi = x.y;
        const double2* r = Rs + c * QW_MAX + q0;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          const double2 y = r[t];
          sr[t] = fma(xr, y.x, fma(xi, y.y, sr[t]));
          si[t] = fma(xi, y.x, fma(-xr, y.y, si[t]));
        }
      }
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      if (q0 + t < k) {
        S2 o;
        o.x = sr[t];
        o.y = si[t];
        if (ab.side == 0)
          ab.Bo[(size_t)p * ab.ldbo + q0 + t] = o;
        else
          ab.Bo[(size_t)(q0 + t) * ab.ldbo + p] = o;
      }
    }
  }
}

template <typename S2, int RL>
__global__ __launch_bounds__(RL == 4 ? 1024 : 768) void qr_reg_kernel(const S2* __restrict__ A, int64_t lda, int m,
                                                                      int n, S2* __restrict__ Qo, int64_t ldq,
                                                                      S2* __restrict__ Ro, int64_t ldr, int flags,
                                                                      const QwAbsorb<S2> ab) {
  __shared__ double2 V[QW_MAX][QW_MAX];    // V[j][(i % RL) * (64 / RL) + i / RL] = v_j(i)
  __shared__ double2 T[QW_MAX];            // tau_j
  __shared__ double2 Rs[QW_MAX * QW_MAX];  // R for the absorption
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int ncw = (n + RL - 1) / RL;  // column waves; the QW_PROD waves after them absorb R
  const int cl = lane % RL, rb = lane / RL;
  const int c = RL * wave + cl;  // this lane's column; rows RL rb + r
  const int k = min(m, n);
  const bool in_adj = (flags & TENET_QR_IN_ADJ) != 0;
  double ar[RL], ai[RL];
#pragma unroll
  for (int r = 0; r < RL; ++r) {  // branch-free: lanes outside A read element 0 and drop it
    const int i = RL * rb + r;
    const bool ok = i < m && c < n;
    const S2 x = A[ok ? (in_adj ? (size_t)i * lda + c : (size_t)c * lda + i) : 0];  // A^H: conj(A(c, i))
    ar[r] = ok ? (double)x.x : 0.0;
    ai[r] = ok ? (in_adj ? -(double)x.y : (double)x.y) : 0.0;
  }
  if (wave == 0) qw_reflector<RL>(ar, ai, 0, cl, rb, V, T);
  __syncthreads();
  for (int j = 0; j < k; ++j) {
    if (wave < ncw && RL * wave + RL - 1 > j) {  // a column wave with columns right of j (uniform)
      const double2 t = T[j];
      double vr[RL], vi[RL], dr, di;
      qw_load_v<RL>(V[j], rb, vr, vi);
      qw_dot<RL>(vr, vi, ar, ai, dr, di);
      if (c > j && c < n) qw_axpy<RL>(vr, vi, t.x * dr + t.y * di, t.x * di - t.y * dr, ar, ai);  // conj(tau) v^H a
      if (j + 1 < k && wave == (j + 1) / RL) qw_reflector<RL>(ar, ai, j + 1, cl, rb, V, T);
    }
    __syncthreads();
  }
  // R (k x n): rows i < k of the registers (zero below the diagonal)
  const bool qadj = (flags & TENET_QR_Q_ADJ) != 0, radj = (flags & TENET_QR_R_ADJ) != 0;
  if (ab.B != nullptr) {  // uniform: R to LDS, then the product waves absorb it
    if (c < n) {
#pragma unroll
      for (int r = 0; r < RL; ++r) {
        double2 x;
        x.x = ar[r];
        x.y = ai[r];
        if (RL * rb + r < k) Rs[c * QW_MAX + RL * rb + r] = x;
      }
    }
    __syncthreads();
    if (wave >= ncw) {
      qw_absorb(ab, Rs, n, k, tid - 64 * ncw);
      return;
    }
  }
  if (Ro != nullptr && c < n) {
#pragma unroll
    for (int r = 0; r < RL; ++r) {
      const int i = RL * rb + r;
      if (i < k) {
        S2 o;
        o.x = ar[r];
        o.y = radj ? -ai[r] : ai[r];
        if (radj)
          Ro[(size_t)i * ldr + c] = o;
        else
          Ro[(size_t)c * ldr + i] = o;
      }
    }
  }
  // Q = H_1 ... H_k [I; 0], backwards; a column left of j is still e_c and
  // v_j vanishes above row j, so H_j leaves it alone.  The next reflector is
  // read while the current one is applied.
  if (RL * wave >= k) return;  // no barrier follows
  double qr[RL], qi[RL];
#pragma unroll
  for (int r = 0; r < RL; ++r) {
    qr[r] = (RL * rb + r == c) ? 1.0 : 0.0;
    qi[r] = 0.0;
  }
  int j = min(k - 1, RL * wave + RL - 1);
  double vr[RL], vi[RL];
  qw_load_v<RL>(V[j], rb, vr, vi);
  for (; j >= 0; --j) {
    const double2 t = T[j];
    double nr[RL], ni[RL], dr, di;
    qw_load_v<RL>(V[j > 0 ? j - 1 : 0], rb, nr, ni);
    qw_dot<RL>(vr, vi, qr, qi, dr, di);
    qw_axpy<RL>(vr, vi, t.x * dr - t.y * di, t.x * di + t.y * dr, qr, qi);  // tau v^H q_c
#pragma unroll
    for (int r = 0; r < RL; ++r) {
      vr[r] = nr[r];
      vi[r] = ni[r];
    }
  }
  if (c < k) {
#pragma unroll
    for (int r = 0; r < RL; ++r) {
      const int i = RL * rb + r;
      if (i < m) {
        S2 o;
        o.x = qr[r];
        o.y = qadj ? -qi[r] : qi[r];
        if (qadj)
          Qo[(size_t)i * ldq + c] = o;
        else
          Qo[(size_t)c * ldq + i] = o;
      }
    }
  }
}

bool qr_small_fits(int64_t m, int64_t n) { return m > 0 && n > 0 && m <= QW_MAX && n <= QW_MAX; }

template <typename S2>
static void launch_qr_reg(const void* A, int64_t lda, int64_t m, int64_t n, void* Qout, int64_t ldq, void* Rout,
                          int64_t ldr, int flags, const QrAbsorbArgs* abs, hipStream_t s, hipEvent_t ev0,
                          hipEvent_t ev1) {
  QwAbsorb<S2> ab{nullptr, nullptr, 0, 0, 0, 0};
  if (abs != nullptr) ab = QwAbsorb<S2>{(const S2*)abs->B, (S2*)abs->Bout, abs->ldb, abs->ldbo, (int)abs->nb, abs->side};
  const int prod = abs != nullptr ? QW_PROD : 0;
  const int w4 = (int)((n + 3) / 4) + prod;
  if (w4 <= 16) {  // four rows per lane
    hipExtLaunchKernelGGL(qr_reg_kernel<S2, 4>, dim3(1), dim3(64 * w4), 0, s, ev0, ev1, 0, (const S2*)A, lda, (int)m,
                          (int)n, (S2*)Qout, ldq, (S2*)Rout, ldr, flags, ab);
  } else {
    const int w8 = (int)((n + 7) / 8) + prod;
    hipExtLaunchKernelGGL(qr_reg_kernel<S2, 8>, dim3(1), dim3(64 * w8), 0, s, ev0, ev1, 0, (const S2*)A, lda, (int)m,
                          (int)n, (S2*)Qout, ldq, (S2*)Rout, ldr, flags, ab);
  }
}

// ev0 / ev1 (optional): stamped at the kernel's start and end by the dispatch
int launch_qr_small(const void* A, int64_t lda, int64_t m, int64_t n, void* Qout, int64_t ldq, void* Rout,
                    int64_t ldr, int dtype, int flags, const QrAbsorbArgs* abs, hipStream_t s, hipEvent_t ev0,
                    hipEvent_t ev1) {
  if (!qr_small_fits(m, n)) {
    set_error("tenet_qr_thin: matrix too large for the single-workgroup QR");
    return TENET_ERR_UNSUPPORTED;
  }
  if (dtype == TENET_C128)
    launch_qr_reg<double2>(A, lda, m, n, Qout, ldq, Rout, ldr, flags, abs, s, ev0, ev1);
  else
    launch_qr_reg<float2>(A, lda, m, n, Qout, ldq, Rout, ldr, flags, abs, s, ev0, ev1);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

// W (m x n, FP64 complex, already packed) -> Q, R outputs.  Workspace after W:
// Q scratch (m x k), tau (k), T blocks.
int launch_qr_blocked(double* W, int64_t m, int64_t n, void* Qout, int64_t ldq, void* Rout, int64_t ldr, int dtype,
                      int flags, hipStream_t s) {
  const int64_t k = std::min(m, n);
  const int64_t np = (k + QB - 1) / QB;
  double* Q = W + (size_t)m * n * 2;
  double* tau = Q + (size_t)m * k * 2;
  double* Tb = tau + (size_t)k * 2;
  double* Yp = Tb + (size_t)np * QB * QB * 2;  // chunk partials of V^H M
  for (int64_t p = 0; p < np; ++p) {
    const int j0 = (int)(p * QB), jb = (int)std::min<int64_t>(QB, k - j0);
    double* Tp = Tb + (size_t)p * QB * QB * 2;
    const int mr = (int)m - j0;
    if (mr <= 64)
      hipLaunchKernelGGL(qr_panel_v4_kernel<1>, dim3(1), dim3(64 * QB), 0, s, W, tau, Tp, (int)m, j0, jb);
    else if (mr <= 128)
      hipLaunchKernelGGL(qr_panel_v4_kernel<2>, dim3(1), dim3(64 * QB), 0, s, W, tau, Tp, (int)m, j0, jb);
    else if (mr <= 256)
      hipLaunchKernelGGL(qr_panel_v4_kernel<4>, dim3(1), dim3(64 * QB), 0, s, W, tau, Tp, (int)m, j0, jb);
    else if (mr <= 512)
      hipLaunchKernelGGL(qr_panel_v4_kernel<8>, dim3(1), dim3(64 * QB), 0, s, W, tau, Tp, (int)m, j0, jb);
    else
      hipLaunchKernelGGL(qr_panel_v3_kernel<4>, dim3(1), dim3(QP3_THREADS), 0, s, W, tau, Tb + (size_t)p * QB * QB * 2,
                         (int)m, j0, jb);
    TENET_HIP_CHECK(hipGetLastError());
    const int cb = j0 + jb;
    if (cb < n) {
      const int st = qr_apply_chunked(W, Tb + (size_t)p * QB * QB * 2, W, (int)m, (int)m, j0, jb, cb, (int)n, 1, Yp, s);
      if (st != TENET_OK) return st;
    }
  }
  {
    const size_t total = (size_t)m * k;
    const unsigned g = (unsigned)std::max<size_t>(1, std::min<size_t>((total + 255) / 256, 4096));
    hipLaunchKernelGGL(qr_eye_kernel, dim3(g), dim3(256), 0, s, Q, (int)m, (int)k);
    TENET_HIP_CHECK(hipGetLastError());
  }
  for (int64_t p = np - 1; p >= 0; --p) {
    const int j0 = (int)(p * QB), jb = (int)std::min<int64_t>(QB, k - j0);
    const int st = qr_apply_chunked(W, Tb + (size_t)p * QB * QB * 2, Q, (int)m, (int)m, j0, jb, j0, (int)k, 0, Yp, s);
    if (st != TENET_OK) return st;
  }
  const size_t total = (size_t)m * k + (size_t)k * n;
  const unsigned g = (unsigned)std::max<size_t>(1, std::min<size_t>((total + 255) / 256, 4096));
  if (dtype == TENET_C128)
    hipLaunchKernelGGL(qr_store_kernel<double2>, dim3(g), dim3(256), 0, s, W, Q, (int)m, (int)n, Qout, ldq, Rout, ldr,
                       flags);
  else
    hipLaunchKernelGGL(qr_store_kernel<float2>, dim3(g), dim3(256), 0, s, W, Q, (int)m, (int)n, Qout, ldq, Rout, ldr,
                       flags);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

}  // namespace tenet
