// K3: thin SVD by one-sided (Hestenes) block Jacobi in FP64, plus the
// truncation rule of Tenet's compress!/simple_update (compress.jl:46-59).
//
// The columns of the work matrix X (nr x ncols, nr >= ncols; a wide input is
// processed as A^H) are cut into blocks of b columns.  A sweep visits every
// block pair once in the round-robin ("circle") order, so the nblk/2 pairs
// of one round are independent.  For a pair (I, J) the 2b x 2b Gram matrix
// G = [X_I X_J]^H [X_I X_J] is diagonalised by cyclic two-sided Jacobi
// (Rutishauser's rotation formulas, complex form: Demmel & Veselic, SIAM J.
// Matrix Anal. Appl. 13 (1992)) in LDS, and the accumulated unitary W is
// applied to the pair's columns of X and V.  At convergence X = U diag(s),
// A = X V^H.  All rotations, Gram sums and norms are FP64.
//
// Each pair may be split over q workgroups by rows (phase 0: partial Gram
// per row chunk; phase 1: every workgroup of the pair sums the q partials in
// a fixed order, solves the same 2b x 2b problem, and rotates its own row
// chunk).  With q = 1 the two phases fuse (phase 2).
#include "../../include/tenet_hip.h"
#include "internal.h"

namespace tenet {

constexpr int SV_THREADS = 256;
constexpr int B2MAX = 32;  // 2b
constexpr int RS = 64;     // staged rows per LDS pass
constexpr int SORT_MAX = 4096;

__device__ __forceinline__ int circle_pos(int i, int r, int nb) {
  return i == 0 ? 0 : 1 + ((i - 1 + r) % (nb - 1));
}

struct SvSmem {
  double gr[B2MAX][B2MAX + 1];
  double gi[B2MAX][B2MAX + 1];
  double wr[B2MAX][B2MAX + 1];
  double wi[B2MAX][B2MAX + 1];
  double xr[RS][B2MAX + 1];
  double xi[RS][B2MAX + 1];
  double rc[B2MAX / 2], rsr[B2MAX / 2], rsi[B2MAX / 2];
  int rp[B2MAX / 2], rq[B2MAX / 2], ract[B2MAX / 2];
  int any_rot, any_outer;
};

__device__ __forceinline__ int pair_col(int l, int I, int J, int b) { return l < b ? I * b + l : J * b + (l - b); }

// Stage rows [r0, r0+rows) of the pair's 2b columns of M (ld) into LDS.
__device__ __forceinline__ void stage_rows(SvSmem& sm, const double* M, int64_t ld, int r0, int rows, int I, int J,
                                           int b) {
  const int n2 = 2 * b;
  for (int e = threadIdx.x; e < RS * n2; e += SV_THREADS) {
    const int r = e % RS, l = e / RS;
    double vr = 0.0, vi = 0.0;
    if (r < rows) {
      const double2 v = reinterpret_cast<const double2*>(M)[(size_t)pair_col(l, I, J, b) * ld + r0 + r];
      vr = v.x;
      vi = v.y;
    }
    sm.xr[r][l] = vr;
    sm.xi[r][l] = vi;
  }
}

// acc (this thread's Gram entries) += staged^H staged
constexpr int GE_PER_THREAD = (B2MAX * B2MAX + SV_THREADS - 1) / SV_THREADS;

__device__ __forceinline__ void gram_accum(SvSmem& sm, int n2, int rows, double* ar, double* ai) {
#pragma unroll
  for (int t = 0; t < GE_PER_THREAD; ++t) {
    const int e = threadIdx.x + t * SV_THREADS;
    if (e >= n2 * n2) continue;
    const int l1 = e % n2, l2 = e / n2;
    double sr = 0.0, si = 0.0;
    for (int r = 0; r < rows; ++r) {
      const double x1r = sm.xr[r][l1], x1i = sm.xi[r][l1];
      const double x2r = sm.xr[r][l2], x2i = sm.xi[r][l2];
      sr += x1r * x2r + x1i * x2i;  // conj(x1) * x2
      si += x1r * x2i - x1i * x2r;
    }
    ar[t] += sr;
    ai[t] += si;
  }
}

__device__ void gram_over_rows(SvSmem& sm, const double* X, int64_t ld, int r0, int r1, int I, int J, int b,
                               double* ar, double* ai) {
  const int n2 = 2 * b;
  for (int t = 0; t < GE_PER_THREAD; ++t) {
    ar[t] = 0.0;
    ai[t] = 0.0;
  }
  for (int rb = r0; rb < r1; rb += RS) {
    const int rows = min(RS, r1 - rb);
    __syncthreads();
    stage_rows(sm, X, ld, rb, rows, I, J, b);
    __syncthreads();
    gram_accum(sm, n2, rows, ar, ai);
  }
}

// Diagonalise the Hermitian sm.g (n2 x n2) in place, W accumulates the
// rotations.  Returns (via sm.any_outer) whether the first sweep found any
// off-diagonal entry above tol (the pair was not yet orthogonal).
__device__ void jacobi_inner(SvSmem& sm, int n2, double tol, int max_sweeps) {
  const int tid = threadIdx.x;
  for (int e = tid; e < n2 * n2; e += SV_THREADS) {
    const int i = e % n2, j = e / n2;
    sm.wr[i][j] = (i == j) ? 1.0 : 0.0;
    sm.wi[i][j] = 0.0;
  }
  if (tid == 0) sm.any_outer = 0;
  __syncthreads();
  const int npair = n2 / 2;
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    if (tid == 0) sm.any_rot = 0;
    __syncthreads();
    for (int step = 0; step < n2 - 1; ++step) {
      if (tid < npair) {
        int p = circle_pos(tid, step, n2), q = circle_pos(n2 - 1 - tid, step, n2);
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        const double a = sm.gr[p][p], d = sm.gr[q][q];
        const double gr = sm.gr[p][q], gi = sm.gi[p][q];
        const double g = sqrt(gr * gr + gi * gi);
        int act = 0;
        double c = 1.0, sr = 0.0, si = 0.0;
        if (g > tol * sqrt(fmax(a, 0.0) * fmax(d, 0.0)) && g > 1e-300) {
          act = 1;
          const double tau = (d - a) / (2.0 * g);
          const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
          c = 1.0 / sqrt(1.0 + t * t);
          const double s = t * c;
          // s * e^{i phi}, e^{i phi} = g_pq / |g_pq|
          sr = s * gr / g;
          si = s * gi / g;
        }
        sm.rp[tid] = p;
        sm.rq[tid] = q;
        sm.rc[tid] = c;
        sm.rsr[tid] = sr;
        sm.rsi[tid] = si;
        sm.ract[tid] = act;
        if (act) {
          sm.any_rot = 1;
          if (sweep == 0) sm.any_outer = 1;
        }
      }
      __syncthreads();
      // J = [[c, s e^{i phi}], [-s e^{-i phi}, c]] on (p, q).
      // rows:  G_p <- c G_p - s e^{i phi} G_q ;  G_q <- s e^{-i phi} G_p + c G_q
      for (int e = tid; e < npair * n2; e += SV_THREADS) {
        const int k = e / n2, col = e % n2;
        if (!sm.ract[k]) continue;
        const int p = sm.rp[k], q = sm.rq[k];
        const double c = sm.rc[k], sr = sm.rsr[k], si = sm.rsi[k];
        const double pr = sm.gr[p][col], pi = sm.gi[p][col];
        const double qr = sm.gr[q][col], qi = sm.gi[q][col];
        sm.gr[p][col] = c * pr - (sr * qr - si * qi);
        sm.gi[p][col] = c * pi - (sr * qi + si * qr);
        sm.gr[q][col] = (sr * pr + si * pi) + c * qr;
        sm.gi[q][col] = (sr * pi - si * pr) + c * qi;
      }
      __syncthreads();
      // columns of G and W:  M_p <- c M_p - s e^{-i phi} M_q ; M_q <- s e^{i phi} M_p + c M_q
      for (int e = tid; e < 2 * npair * n2; e += SV_THREADS) {
        const int which = e / (npair * n2);
        const int ee = e % (npair * n2);
        const int k = ee / n2, row = ee % n2;
        if (!sm.ract[k]) continue;
        const int p = sm.rp[k], q = sm.rq[k];
        const double c = sm.rc[k], sr = sm.rsr[k], si = sm.rsi[k];
        double(*Mr)[B2MAX + 1] = which ? sm.wr : sm.gr;
        double(*Mi)[B2MAX + 1] = which ? sm.wi : sm.gi;
        const double pr = Mr[row][p], pi = Mi[row][p];
        const double qr = Mr[row][q], qi = Mi[row][q];
        Mr[row][p] = c * pr - (sr * qr + si * qi);
        Mi[row][p] = c * pi - (sr * qi - si * qr);
        Mr[row][q] = (sr * pr - si * pi) + c * qr;
        Mi[row][q] = (sr * pi + si * pr) + c * qi;
      }
      __syncthreads();
      if (tid < npair && sm.ract[tid]) {
        const int p = sm.rp[tid], q = sm.rq[tid];
        sm.gr[p][q] = 0.0;
        sm.gi[p][q] = 0.0;
        sm.gr[q][p] = 0.0;
        sm.gi[q][p] = 0.0;
        sm.gi[p][p] = 0.0;
        sm.gi[q][q] = 0.0;
      }
      __syncthreads();
    }
    const int rotated = sm.any_rot;
    __syncthreads();  // everyone has read any_rot before thread 0 resets it
    if (!rotated) break;
  }
}

// rows [r0, r1) of the pair's columns of M <- M W
__device__ void apply_rows(SvSmem& sm, double* M, int64_t ld, int r0, int r1, int I, int J, int b) {
  const int n2 = 2 * b;
  for (int rb = r0; rb < r1; rb += RS) {
    const int rows = min(RS, r1 - rb);
    __syncthreads();
    stage_rows(sm, M, ld, rb, rows, I, J, b);
    __syncthreads();
    for (int e = threadIdx.x; e < RS * n2; e += SV_THREADS) {
      const int r = e % RS, l = e / RS;
      if (r >= rows) continue;
      double sr = 0.0, si = 0.0;
      for (int t = 0; t < n2; ++t) {
        const double xr = sm.xr[r][t], xi = sm.xi[r][t];
        const double wr = sm.wr[t][l], wi = sm.wi[t][l];
        sr += xr * wr - xi * wi;
        si += xr * wi + xi * wr;
      }
      double2 v;
      v.x = sr;
      v.y = si;
      reinterpret_cast<double2*>(M)[(size_t)pair_col(l, I, J, b) * ld + rb + r] = v;
    }
  }
  __syncthreads();
}

__global__ __launch_bounds__(SV_THREADS) void svd_round_kernel(SvdBatch Bt, int round, int phase) {
  __shared__ SvSmem sm;
  const SvdProblem& P = Bt.arr ? Bt.arr[blockIdx.y] : Bt.one;
  if (P.flag[1]) return;  // converged
  const int nb = P.nblk;
  if (round >= nb - 1) return;
  const int q = P.q;
  const int item = blockIdx.x;
  const int p = item / q, c = item % q;
  if (p >= nb / 2) return;
  int I = circle_pos(p, round, nb), J = circle_pos(nb - 1 - p, round, nb);
  if (I > J) {
    const int t = I;
    I = J;
    J = t;
  }
  const int b = P.b, n2 = 2 * b;
  const int rcx = (P.nr + q - 1) / q;
  const int r0 = c * rcx, r1 = min(P.nr, r0 + rcx);
  double ar[GE_PER_THREAD], ai[GE_PER_THREAD];

  if (phase == 0) {
    gram_over_rows(sm, P.X, P.nr, r0, r1, I, J, b, ar, ai);
    double* out = P.Gpart + ((size_t)p * q + c) * (size_t)n2 * n2 * 2;
#pragma unroll
    for (int t = 0; t < GE_PER_THREAD; ++t) {
      const int e = threadIdx.x + t * SV_THREADS;
      if (e >= n2 * n2) continue;
      out[2 * e] = ar[t];
      out[2 * e + 1] = ai[t];
    }
    return;
  }
  if (phase == 2) {
    gram_over_rows(sm, P.X, P.nr, 0, P.nr, I, J, b, ar, ai);
  } else {
    for (int t = 0; t < GE_PER_THREAD; ++t) {
      ar[t] = 0.0;
      ai[t] = 0.0;
    }
    for (int cc = 0; cc < q; ++cc) {
      const double* in = P.Gpart + ((size_t)p * q + cc) * (size_t)n2 * n2 * 2;
#pragma unroll
      for (int t = 0; t < GE_PER_THREAD; ++t) {
        const int e = threadIdx.x + t * SV_THREADS;
        if (e >= n2 * n2) continue;
        ar[t] += in[2 * e];
        ai[t] += in[2 * e + 1];
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < GE_PER_THREAD; ++t) {
    const int e = threadIdx.x + t * SV_THREADS;
    if (e >= n2 * n2) continue;
    const int l1 = e % n2, l2 = e / n2;
    sm.gr[l1][l2] = ar[t];
    sm.gi[l1][l2] = ai[t];
  }
  __syncthreads();
  jacobi_inner(sm, n2, P.tol, 6);
  if (!sm.any_outer) return;  // pair already orthogonal: W = I (uniform branch)
  if (threadIdx.x == 0 && c == 0) P.flag[0] = 1;
  const int rx0 = (phase == 2) ? 0 : r0, rx1 = (phase == 2) ? P.nr : r1;
  apply_rows(sm, P.X, P.nr, rx0, rx1, I, J, b);
  const int rcv = (P.ncp + q - 1) / q;
  const int v0 = (phase == 2) ? 0 : c * rcv;
  const int v1 = (phase == 2) ? P.ncp : min(P.ncp, v0 + rcv);
  apply_rows(sm, P.V, P.ncp, v0, v1, I, J, b);
}

__global__ void svd_sweep_end_kernel(SvdBatch Bt, int64_t nprob) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nprob) return;
  const SvdProblem& P = Bt.arr ? Bt.arr[i] : Bt.one;
  if (P.flag[1]) return;
  if (P.flag[0] == 0) P.flag[1] = 1;
  P.flag[0] = 0;
  P.flag[2] += 1;
}

// X (nr x ncp) from A or A^H with zero padding; V = I.
template <typename S2>
__global__ void svd_init_kernel(SvdBatch Bt) {
  const SvdProblem& P = Bt.arr ? Bt.arr[blockIdx.y] : Bt.one;
  const size_t nx = (size_t)P.nr * P.ncp, nv = (size_t)P.ncp * P.ncp;
  const S2* A = static_cast<const S2*>(P.A);
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nx + nv; e += (size_t)gridDim.x * blockDim.x) {
    double2 v;
    v.x = 0.0;
    v.y = 0.0;
    if (e < nx) {
      const int r = (int)(e % P.nr), c = (int)(e / P.nr);
      if (c < P.ncols) {
        if (!P.transposed) {
          const S2 a = A[(size_t)c * P.lda + r];
          v.x = a.x;
          v.y = a.y;
        } else {  // X = A^H : X[r][c] = conj(A[c][r])
          const S2 a = A[(size_t)r * P.lda + c];
          v.x = a.x;
          v.y = -(double)a.y;
        }
      }
      reinterpret_cast<double2*>(P.X)[e] = v;
    } else {
      const size_t f = e - nx;
      const int r = (int)(f % P.ncp), c = (int)(f / P.ncp);
      v.x = (r == c) ? 1.0 : 0.0;
      reinterpret_cast<double2*>(P.V)[f] = v;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    P.flag[0] = 0;
    P.flag[1] = 0;
    P.flag[2] = 0;
  }
}

// norms, descending sort, S, kept count and discarded weight
__global__ __launch_bounds__(SV_THREADS) void svd_sort_kernel(SvdBatch Bt) {
  __shared__ double key[SORT_MAX];
  __shared__ int idx[SORT_MAX];
  __shared__ double red[SV_THREADS / 64];
  __shared__ int cnt[SV_THREADS / 64];
  const SvdProblem& P = Bt.arr ? Bt.arr[blockIdx.x] : Bt.one;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int ncp = P.ncp;
  int np2 = 1;
  while (np2 < ncp) np2 <<= 1;
  for (int j = wave; j < np2; j += SV_THREADS / 64) {
    double s = 0.0;
    if (j < ncp) {
      const double2* col = reinterpret_cast<const double2*>(P.X) + (size_t)j * P.nr;
      for (int r = lane; r < P.nr; r += 64) {
        const double2 v = col[r];
        s += v.x * v.x + v.y * v.y;
      }
      for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    }
    if (lane == 0) {
      key[j] = (j < ncp) ? sqrt(s) : -1.0;
      idx[j] = j;
    }
  }
  __syncthreads();
  // bitonic sort, descending by key, ties by index
  for (int size = 2; size <= np2; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = tid; i < np2; i += SV_THREADS) {
        const int j = i ^ stride;
        if (j > i) {
          const bool desc = ((i & size) == 0);
          const double ki = key[i], kj = key[j];
          const int ii = idx[i], ij = idx[j];
          const bool i_first = (ki > kj) || (ki == kj && ii < ij);
          if (desc != i_first) {
            key[i] = kj;
            key[j] = ki;
            idx[i] = ij;
            idx[j] = ii;
          }
        }
      }
      __syncthreads();
    }
  }
  const int kmax = P.ncols;
  for (int j = tid; j < kmax; j += SV_THREADS) {
    P.sigma[j] = key[j];
    P.perm[j] = idx[j];
    if (P.S) P.S[j] = key[j];
  }
  // kept count: min(maxdim, #(s > threshold)) -- a prefix since s descends
  int kcap = (P.maxdim > 0 && P.maxdim < kmax) ? (int)P.maxdim : kmax;
  int c = 0;
  for (int j = tid; j < kcap; j += SV_THREADS) c += (P.threshold > 0.0) ? (key[j] > P.threshold) : 1;
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
  if (lane == 0) cnt[wave] = c;
  __syncthreads();
  int k = 0;
  for (int w = 0; w < SV_THREADS / 64; ++w) k += cnt[w];
  double d = 0.0;
  for (int j = k + tid; j < kmax; j += SV_THREADS) d += key[j] * key[j];
  for (int off = 32; off > 0; off >>= 1) d += __shfl_xor(d, off, 64);
  if (lane == 0) red[wave] = d;
  __syncthreads();
  if (tid == 0) {
    double disc = 0.0;
    for (int w = 0; w < SV_THREADS / 64; ++w) disc += red[w];
    if (P.k_dev) *P.k_dev = k;
    if (P.disc_dev) *P.disc_dev = disc;
  }
}

// U (m x kmax) and Vh (kmax x n) from X, V and the permutation.
template <typename S2>
__global__ void svd_write_kernel(SvdBatch Bt) {
  const SvdProblem& P = Bt.arr ? Bt.arr[blockIdx.y] : Bt.one;
  const int kmax = P.ncols;
  const size_t nu = (size_t)P.m * kmax, nvh = (size_t)kmax * P.n;
  const double2* X = reinterpret_cast<const double2*>(P.X);
  const double2* V = reinterpret_cast<const double2*>(P.V);
  S2* U = static_cast<S2*>(P.U);
  S2* Vh = static_cast<S2*>(P.Vh);
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nu + nvh; e += (size_t)gridDim.x * blockDim.x) {
    S2 o;
    if (e < nu) {
      const int r = (int)(e % P.m), j = (int)(e / P.m);
      const int pj = P.perm[j];
      const double s = P.sigma[j];
      double vr, vi;
      if (!P.transposed) {
        const double2 x = X[(size_t)pj * P.nr + r];
        vr = s > 0.0 ? x.x / s : 0.0;
        vi = s > 0.0 ? x.y / s : 0.0;
      } else {
        const double2 x = V[(size_t)pj * P.ncp + r];
        vr = x.x;
        vi = x.y;
      }
      o.x = vr;
      o.y = vi;
      U[(size_t)j * P.ldu + r] = o;
    } else {
      const size_t f = e - nu;
      const int j = (int)(f % kmax), cidx = (int)(f / kmax);
      const int pj = P.perm[j];
      const double s = P.sigma[j];
      double vr, vi;
      if (!P.transposed) {
        const double2 x = V[(size_t)pj * P.ncp + cidx];
        vr = x.x;
        vi = -x.y;
      } else {
        const double2 x = X[(size_t)pj * P.nr + cidx];
        vr = s > 0.0 ? x.x / s : 0.0;
        vi = s > 0.0 ? -x.y / s : 0.0;
      }
      o.x = vr;
      o.y = vi;
      Vh[(size_t)cidx * P.ldvh + j] = o;
    }
  }
}

int launch_svd_init(const SvdBatch& b, int64_t nprob, int64_t max_elems, int dtype, hipStream_t s) {
  const unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>((max_elems + 255) / 256, 1024));
  if (dtype == TENET_C128)
    hipLaunchKernelGGL(svd_init_kernel<double2>, dim3(gx, (unsigned)nprob), dim3(256), 0, s, b);
  else
    hipLaunchKernelGGL(svd_init_kernel<float2>, dim3(gx, (unsigned)nprob), dim3(256), 0, s, b);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

int launch_svd_round(const SvdBatch& b, int64_t nprob, int max_items, int round, int phase, hipStream_t s) {
  hipLaunchKernelGGL(svd_round_kernel, dim3((unsigned)max_items, (unsigned)nprob), dim3(SV_THREADS), 0, s, b, round,
                     phase);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

int launch_svd_sweep_end(const SvdBatch& b, int64_t nprob, hipStream_t s) {
  hipLaunchKernelGGL(svd_sweep_end_kernel, dim3((unsigned)((nprob + 255) / 256)), dim3(256), 0, s, b, nprob);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

int launch_svd_finalize(const SvdBatch& b, int64_t nprob, int max_ncp, int64_t max_elems, int dtype,
                        hipStream_t s) {
  if (max_ncp > SORT_MAX) {
    set_error("SVD: min(m, n) exceeds the sort limit (4096)");
    return TENET_ERR_UNSUPPORTED;
  }
  hipLaunchKernelGGL(svd_sort_kernel, dim3((unsigned)nprob), dim3(SV_THREADS), 0, s, b);
  TENET_HIP_CHECK(hipGetLastError());
  const unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>((max_elems + 255) / 256, 2048));
  if (dtype == TENET_C128)
    hipLaunchKernelGGL(svd_write_kernel<double2>, dim3(gx, (unsigned)nprob), dim3(256), 0, s, b);
  else
    hipLaunchKernelGGL(svd_write_kernel<float2>, dim3(gx, (unsigned)nprob), dim3(256), 0, s, b);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

}  // namespace tenet
