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
#include "../../include/tenet_hip.h"
#include "internal.h"

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
