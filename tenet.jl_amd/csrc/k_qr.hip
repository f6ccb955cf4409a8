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
// pairing lanes l ^ 15, l ^ 7, l ^ 2, l ^ 1 by DPP, l ^ 16 by swizzle, l ^ 32
// by permute; lane l ends with the wave total of value 16 b3 + 8 b2 + 4 b1 +
// 2 b0 + b4, b_i the bits of l), the four wave totals meet in LDS (double
// buffered by column parity): one barrier per column.
template <int CTRL>
__device__ __forceinline__ double dpp_d(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

__device__ __forceinline__ double swz16_d(double x) {  // lane l ^ 16 within each 32-lane half
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_ds_swizzle((int)(b & 0xffffffffll), 0x401F);
  const int hi = __builtin_amdgcn_ds_swizzle((int)(b >> 32), 0x401F);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
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
  {
    const bool hi = (lane & 16) != 0;
    const double keep = hi ? v[1] : v[0];
    const double send = hi ? v[0] : v[1];
    v[0] = keep + swz16_d(send);
  }
  v[0] += __shfl_xor(v[0], 32, 64);
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

__global__ __launch_bounds__(QR_THREADS) void qr_apply_y_kernel(const double* __restrict__ W,
                                                                 const double* __restrict__ M, double* __restrict__ Yp,
                                                                 int m, int ldm, int j0, int jb, int col_begin,
                                                                 int col_end) {
  __shared__ double vr[QA_RC][QB + 1], vi[QA_RC][QB + 1];
  __shared__ double xr[QA_RC][QB + 1], xi[QA_RC][QB + 1];
  __shared__ double yred[4][QB * QB][2];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int blk = blockIdx.x, chunk = blockIdx.y;
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
      double a = yr_[i][j], b = yi_[i][j];
      a += __shfl_xor(a, 16, 64);
      b += __shfl_xor(b, 16, 64);
      a += __shfl_xor(a, 32, 64);
      b += __shfl_xor(b, 32, 64);
      yr_[i][j] = a;
      yi_[i][j] = b;
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
  const int nblk = gridDim.x;
  double* out = Yp + ((size_t)chunk * nblk + blk) * QB * QB * 2;
  {
    const int e = tid;  // QB * QB = 256 = QR_THREADS
    out[2 * e] = yred[0][e][0] + yred[1][e][0] + yred[2][e][0] + yred[3][e][0];
    out[2 * e + 1] = yred[0][e][1] + yred[1][e][1] + yred[2][e][1] + yred[3][e][1];
  }
}

__global__ __launch_bounds__(QR_THREADS) void qr_apply_x_kernel(const double* __restrict__ W,
                                                                 const double* __restrict__ T,
                                                                 const double* __restrict__ Yp, double* __restrict__ M,
                                                                 int m, int ldm, int j0, int jb, int col_begin,
                                                                 int col_end, int nchunks, int adj) {
  __shared__ double vr[QA_RC][QB + 1], vi[QA_RC][QB + 1];
  __shared__ double yr[QB][QB + 1], yi[QB][QB + 1];
  __shared__ double zr[QB][QB + 1], zi[QB][QB + 1];
  const int tid = threadIdx.x;
  const int blk = blockIdx.x, chunk = blockIdx.y;
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
    const int nblk = gridDim.x;
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

// The panel kernel keeps up to 512 x 16 complex doubles (128 KiB) in dynamic
// LDS, above HIP's default dynamic limit; raise it once per process.
size_t qr_blocked_ws_bytes(int64_t m, int64_t n) {
  const int64_t k = std::min(m, n);
  const int64_t np = (k + QB - 1) / QB;
  const int64_t ych = (m + QA_RC - 1) / QA_RC, yblk = (std::max(n, k) + QB - 1) / QB;
  return (size_t)m * n * 16 + (size_t)m * k * 16 + (size_t)k * 16 + (size_t)np * QB * QB * 16 +
         (size_t)ych * yblk * QB * QB * 16 + 2048;
}

// W (m x n, FP64 complex, already packed) -> Q, R outputs.  Workspace after W:
// Q scratch (m x k), tau (k), T blocks.
// ---------------------------------------------------------------------------
// Small QR in one workgroup: A (m x n) and Q (m x k) resident in LDS for
// m, n <= 64 (the canonize! moves at small bond dimensions: 64 x 32,
// 2 x 32, ...), one launch instead of the blocked path's panel / apply /
// Q-build sequence.  The reflectors are those of
// qr_factor_kernel (v(0) = 1, beta = -sign(Re alpha) ||(alpha, x)||, H^H
// applied from the left, Q = H_1 ... H_k [I; 0] accumulated backwards).
// Lanes run over rows and the sixteen waves over columns (two columns per
// pass, so two reductions are in flight).  Wave 0 owns column j + 1 at step j: it
// updates that column first and then forms the next reflector, so a
// reflector costs one barrier; column j's scaled vector is written during
// step j + 1, when no wave reads that column.
// ---------------------------------------------------------------------------
constexpr int QS_THREADS = 1024;
constexpr int QS_ELEMS = 8192;  // complex entries of A and Q together (128 KiB)
constexpr int QS_MAX_K = 64;    // m (n + k) <= QS_ELEMS and k = min(m, n) give k <= 64

// Sum over the 64 lanes, every lane receiving it: quad, half-row and row
// levels by DPP (register latency), then the 32-lane and 64-lane halves.
__device__ __forceinline__ double wave_sum_dpp(double v) {
  v += dpp_d<0xB1>(v);   // quad_perm [1,0,3,2]: l ^ 1
  v += dpp_d<0x4E>(v);   // quad_perm [2,3,0,1]: l ^ 2
  v += dpp_d<0x141>(v);  // row_half_mirror: the other quad of the 8
  v += dpp_d<0x140>(v);  // row_mirror: the other 8 of the 16
  v += swz16_d(v);       // l ^ 16
  v += __shfl_xor(v, 32, 64);
  return v;
}

struct QsScal {
  double tr, ti;  // tau
  double sr, si;  // 1 / (alpha - beta)
  double beta;
  int ident;  // H = I
};

// Reflector of column j (rows j..m-1), one wave; lane 0 writes *out.
__device__ __forceinline__ void qs_reflector(const double2* A, int m, int j, int lane, QsScal* out) {
  double part = 0.0;
  for (int i = j + 1 + lane; i < m; i += 64) {
    const double2 x = A[(size_t)j * m + i];
    part += x.x * x.x + x.y * x.y;
  }
  part = wave_sum_dpp(part);
  if (lane == 0) {
    const double2 a = A[(size_t)j * m + j];
    const double ar = a.x, ai = a.y;
    QsScal o;
    o.tr = 0.0;
    o.ti = 0.0;
    o.sr = 0.0;
    o.si = 0.0;
    o.beta = ar;
    o.ident = 1;
    if (!(part == 0.0 && ai == 0.0)) {
      const double nrm = sqrt(ar * ar + ai * ai + part);
      const double beta = (ar >= 0.0) ? -nrm : nrm;
      o.tr = (beta - ar) / beta;
      o.ti = -ai / beta;
      const double dr = ar - beta, di = ai;
      const double den = dr * dr + di * di;
      o.sr = dr / den;
      o.si = -di / den;
      o.beta = beta;
      o.ident = 0;
    }
    *out = o;
  }
}

template <typename S2>
__global__ __launch_bounds__(QS_THREADS) void qr_small_kernel(const double* __restrict__ W, int m, int n, void* Qout,
                                                              int64_t ldq, void* Rout, int64_t ldr, int flags) {
  __shared__ double2 A[QS_ELEMS];  // A (ld = m), then Q (ld = m)
  __shared__ QsScal sc[QS_MAX_K];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int k = min(m, n);
  double2* Q = A + (size_t)m * n;
  const double2* W2 = reinterpret_cast<const double2*>(W);
  for (int e = tid; e < m * n; e += QS_THREADS) A[e] = W2[e];
  __syncthreads();
  if (wave == 0) qs_reflector(A, m, 0, lane, &sc[0]);
  __syncthreads();
  for (int j = 0; j < k; ++j) {
    if (j > 0 && !sc[j - 1].ident) {
      // column j - 1: beta on the diagonal, the scaled vector below
      const QsScal p = sc[j - 1];
      for (int i = j - 1 + tid; i < m; i += QS_THREADS) {
        double2 x = A[(size_t)(j - 1) * m + i];
        if (i == j - 1) {
          x.x = p.beta;
          x.y = 0.0;
        } else {
          const double xr = x.x, xi = x.y;
          x.x = xr * p.sr - xi * p.si;
          x.y = xr * p.si + xi * p.sr;
        }
        A[(size_t)(j - 1) * m + i] = x;
      }
    }
    const QsScal r = sc[j];
    bool next_done = false;
    if (!r.ident) {
      // A(j:, c) -= conj(tau) v (v^H A(j:, c)), v = [1; x / (alpha - beta)]
      for (int c = j + 1 + wave; c < n; c += 2 * (QS_THREADS / 64)) {
        const int c2 = c + QS_THREADS / 64;
        const bool two = c2 < n;
        double d1r = 0.0, d1i = 0.0, d2r = 0.0, d2i = 0.0;
        for (int i = j + lane; i < m; i += 64) {
          double vr = 1.0, vi = 0.0;
          if (i != j) {
            const double2 x = A[(size_t)j * m + i];
            vr = x.x * r.sr - x.y * r.si;
            vi = x.x * r.si + x.y * r.sr;
          }
          const double2 a1 = A[(size_t)c * m + i];
          d1r += vr * a1.x + vi * a1.y;
          d1i += vr * a1.y - vi * a1.x;
          if (two) {
            const double2 a2 = A[(size_t)c2 * m + i];
            d2r += vr * a2.x + vi * a2.y;
            d2i += vr * a2.y - vi * a2.x;
          }
        }
        d1r = wave_sum_dpp(d1r);
        d1i = wave_sum_dpp(d1i);
        d2r = wave_sum_dpp(d2r);
        d2i = wave_sum_dpp(d2i);
        const double f1r = r.tr * d1r + r.ti * d1i, f1i = r.tr * d1i - r.ti * d1r;  // conj(tau) * dot
        const double f2r = r.tr * d2r + r.ti * d2i, f2i = r.tr * d2i - r.ti * d2r;
        for (int i = j + lane; i < m; i += 64) {
          double vr = 1.0, vi = 0.0;
          if (i != j) {
            const double2 x = A[(size_t)j * m + i];
            vr = x.x * r.sr - x.y * r.si;
            vi = x.x * r.si + x.y * r.sr;
          }
          double2 a1 = A[(size_t)c * m + i];
          a1.x -= vr * f1r - vi * f1i;
          a1.y -= vr * f1i + vi * f1r;
          A[(size_t)c * m + i] = a1;
          if (two) {
            double2 a2 = A[(size_t)c2 * m + i];
            a2.x -= vr * f2r - vi * f2i;
            a2.y -= vr * f2i + vi * f2r;
            A[(size_t)c2 * m + i] = a2;
          }
        }
        if (wave == 0 && !next_done && j + 1 < k) {
          // column j + 1 (this wave's first) is final: the next reflector
          qs_reflector(A, m, j + 1, lane, &sc[j + 1]);
          next_done = true;
        }
      }
    }
    if (wave == 0 && !next_done && j + 1 < k) qs_reflector(A, m, j + 1, lane, &sc[j + 1]);
    __syncthreads();
  }
  if (k > 0 && !sc[k - 1].ident) {
    const QsScal p = sc[k - 1];
    for (int i = k - 1 + tid; i < m; i += QS_THREADS) {
      double2 x = A[(size_t)(k - 1) * m + i];
      if (i == k - 1) {
        x.x = p.beta;
        x.y = 0.0;
      } else {
        const double xr = x.x, xi = x.y;
        x.x = xr * p.sr - xi * p.si;
        x.y = xr * p.si + xi * p.sr;
      }
      A[(size_t)(k - 1) * m + i] = x;
    }
  }
  // Q = H_1 ... H_k [I; 0], backwards
  for (int e = tid; e < m * k; e += QS_THREADS) {
    double2 z;
    z.x = (e % m == e / m) ? 1.0 : 0.0;
    z.y = 0.0;
    Q[e] = z;
  }
  __syncthreads();
  for (int j = k - 1; j >= 0; --j) {
    const QsScal r = sc[j];
    if (r.ident) continue;  // uniform
    // Q(j:, c) -= tau v (v^H Q(j:, c)) for c = j..k-1 (columns < j are still e_c)
    for (int c = j + wave; c < k; c += 2 * (QS_THREADS / 64)) {
      const int c2 = c + QS_THREADS / 64;
      const bool two = c2 < k;
      double d1r = 0.0, d1i = 0.0, d2r = 0.0, d2i = 0.0;
      for (int i = j + lane; i < m; i += 64) {
        double vr = 1.0, vi = 0.0;
        if (i != j) {
          const double2 x = A[(size_t)j * m + i];
          vr = x.x;
          vi = x.y;
        }
        const double2 q1 = Q[(size_t)c * m + i];
        d1r += vr * q1.x + vi * q1.y;
        d1i += vr * q1.y - vi * q1.x;
        if (two) {
          const double2 q2 = Q[(size_t)c2 * m + i];
          d2r += vr * q2.x + vi * q2.y;
          d2i += vr * q2.y - vi * q2.x;
        }
      }
      d1r = wave_sum_dpp(d1r);
      d1i = wave_sum_dpp(d1i);
      d2r = wave_sum_dpp(d2r);
      d2i = wave_sum_dpp(d2i);
      const double f1r = r.tr * d1r - r.ti * d1i, f1i = r.tr * d1i + r.ti * d1r;  // tau * dot
      const double f2r = r.tr * d2r - r.ti * d2i, f2i = r.tr * d2i + r.ti * d2r;
      for (int i = j + lane; i < m; i += 64) {
        double vr = 1.0, vi = 0.0;
        if (i != j) {
          const double2 x = A[(size_t)j * m + i];
          vr = x.x;
          vi = x.y;
        }
        double2 q1 = Q[(size_t)c * m + i];
        q1.x -= vr * f1r - vi * f1i;
        q1.y -= vr * f1i + vi * f1r;
        Q[(size_t)c * m + i] = q1;
        if (two) {
          double2 q2 = Q[(size_t)c2 * m + i];
          q2.x -= vr * f2r - vi * f2i;
          q2.y -= vr * f2i + vi * f2r;
          Q[(size_t)c2 * m + i] = q2;
        }
      }
    }
    __syncthreads();
  }
  // outputs: Q (m x k) and R (k x n, upper triangle of A), adjoints on request
  const bool qadj = (flags & TENET_QR_Q_ADJ) != 0, radj = (flags & TENET_QR_R_ADJ) != 0;
  S2* Qo = static_cast<S2*>(Qout);
  S2* Ro = static_cast<S2*>(Rout);
  for (int e = tid; e < m * k; e += QS_THREADS) {
    const int r = e % m, c = e / m;
    const double2 v = Q[e];
    S2 o;
    o.x = v.x;
    o.y = qadj ? -v.y : v.y;
    if (qadj)
      Qo[(size_t)r * ldq + c] = o;
    else
      Qo[(size_t)c * ldq + r] = o;
  }
  for (int e = tid; e < k * n; e += QS_THREADS) {
    const int r = e % k, c = e / k;
    double2 v;
    v.x = 0.0;
    v.y = 0.0;
    if (r <= c) v = A[(size_t)c * m + r];
    S2 o;
    o.x = v.x;
    o.y = radj ? -v.y : v.y;
    if (radj)
      Ro[(size_t)r * ldr + c] = o;
    else
      Ro[(size_t)c * ldr + r] = o;
  }
}

// one row per lane and at most two column pairs per wave and step; taller or
// wider matrices are faster on the blocked path
bool qr_small_fits(int64_t m, int64_t n) {
  const int64_t k = std::min(m, n);
  return m > 0 && n > 0 && m <= 64 && n <= 64 && m * (n + k) <= QS_ELEMS;
}

// ev0 / ev1 (optional): stamped at the kernel's start and end by the dispatch
int launch_qr_small(const double* W, int64_t m, int64_t n, void* Qout, int64_t ldq, void* Rout, int64_t ldr,
                    int dtype, int flags, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
  if (!qr_small_fits(m, n)) {
    set_error("tenet_qr_thin: matrix too large for the single-workgroup QR");
    return TENET_ERR_UNSUPPORTED;
  }
  if (dtype == TENET_C128)
    hipExtLaunchKernelGGL(qr_small_kernel<double2>, dim3(1), dim3(QS_THREADS), 0, s, ev0, ev1, 0, W, (int)m, (int)n,
                          Qout, ldq, Rout, ldr, flags);
  else
    hipExtLaunchKernelGGL(qr_small_kernel<float2>, dim3(1), dim3(QS_THREADS), 0, s, ev0, ev1, 0, W, (int)m, (int)n,
                          Qout, ldq, Rout, ldr, flags);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

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
    if (m - j0 <= 2 * QP3_THREADS)
      hipLaunchKernelGGL(qr_panel_v3_kernel<2>, dim3(1), dim3(QP3_THREADS), 0, s, W, tau, Tb + (size_t)p * QB * QB * 2,
                         (int)m, j0, jb);
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
