// Small kernels around the hot path: fills, the gate-update epilogue that
// absorbs singular values (Muscle.simple_update's AbsorbEqually,
// simple_update.jl:63), the running log-scale of transfer-matrix sweeps, and
// K4's batched amplitude sweep <x|psi> (overlap.jl:22-32).
#include <algorithm>

#include <hip/hip_ext.h>

#include "../../include/tenet_hip.h"
#include "internal.h"

namespace tenet {

__global__ void fill_zero_kernel(uint32_t* p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0u;
}
__global__ void fill_zero_bytes_kernel(uint8_t* p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0;
}

int launch_fill_zero(void* dst, size_t bytes, hipStream_t s) {
  if (bytes == 0) return TENET_OK;
  if (((uintptr_t)dst % 4) == 0 && bytes % 4 == 0) {
    const size_t n = bytes / 4;
    const unsigned g = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(fill_zero_kernel, dim3(g), dim3(256), 0, s, (uint32_t*)dst, n);
  } else {
    const unsigned g = (unsigned)std::min<size_t>((bytes + 255) / 256, 4096);
    hipLaunchKernelGGL(fill_zero_bytes_kernel, dim3(g), dim3(256), 0, s, (uint8_t*)dst, bytes);
  }
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

// Aout (m x k) = U[:, :k] * fa(S);  Bout (k x n, ld k) = fb(S) * Vh[:k, :] (ld kmax)
template <typename S2>
__global__ void absorb_kernel(const S2* __restrict__ U, int64_t m, const S2* __restrict__ Vh, int64_t ldvh, int64_t n,
                              const double* __restrict__ S, const int64_t* __restrict__ k_dev, int mode,
                              S2* __restrict__ Aout, S2* __restrict__ Bout) {
  const int64_t k = *k_dev;
  const size_t na = (size_t)m * k, nb = (size_t)k * n;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < na + nb; e += (size_t)gridDim.x * blockDim.x) {
    if (e < na) {
      const int64_t j = (int64_t)(e / m);
      const double s = S[j];
      const double f = mode == TENET_ABSORB_EQUAL ? sqrt(s) : (mode == TENET_ABSORB_LEFT ? s : 1.0);
      S2 v = U[e];
      v.x = (double)v.x * f;
      v.y = (double)v.y * f;
      Aout[e] = v;
    } else {
      const size_t f2 = e - na;
      const int64_t j = (int64_t)(f2 % k), c = (int64_t)(f2 / k);
      const double s = S[j];
      const double f = mode == TENET_ABSORB_EQUAL ? sqrt(s) : (mode == TENET_ABSORB_RIGHT ? s : 1.0);
      S2 v = Vh[(size_t)c * ldvh + j];
      v.x = (double)v.x * f;
      v.y = (double)v.y * f;
      Bout[f2] = v;
    }
  }
}

int launch_absorb(const void* U, int64_t m, const void* Vh, int64_t ldvh, int64_t n, const double* S,
                  const int64_t* k_dev, int absorb, void* Aout, void* Bout, int64_t kmax, int dtype, hipStream_t s) {
  const size_t total = (size_t)(m + n) * kmax;
  const unsigned g = (unsigned)std::max<size_t>(1, std::min<size_t>((total + 255) / 256, 2048));
  if (dtype == TENET_C128)
    hipLaunchKernelGGL(absorb_kernel<double2>, dim3(g), dim3(256), 0, s, (const double2*)U, m, (const double2*)Vh,
                       ldvh, n, S, k_dev, absorb, (double2*)Aout, (double2*)Bout);
  else
    hipLaunchKernelGGL(absorb_kernel<float2>, dim3(g), dim3(256), 0, s, (const float2*)U, m, (const float2*)Vh, ldvh,
                       n, S, k_dev, absorb, (float2*)Aout, (float2*)Bout);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ double wave_add(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// x /= max|x| (max over re and im parts); acc[2] += log(max)
template <typename S2>
__global__ __launch_bounds__(1024) void rescale_log_kernel(S2* __restrict__ x, int64_t count, double* acc) {
  __shared__ double red[16];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double mx = 0.0;
  for (int64_t i = tid; i < count; i += blockDim.x) {
    const S2 v = x[i];
    mx = fmax(mx, fmax(fabs((double)v.x), fabs((double)v.y)));
  }
  mx = wave_max(mx);
  if (lane == 0) red[wave] = mx;
  __syncthreads();
  if (tid == 0) {
    double m = 0.0;
    for (int w = 0; w < (int)(blockDim.x / 64); ++w) m = fmax(m, red[w]);
    red[0] = m;
  }
  __syncthreads();
  const double m = red[0];
  if (m > 0.0) {
    const double inv = 1.0 / m;
    for (int64_t i = tid; i < count; i += blockDim.x) {
      S2 v = x[i];
      v.x = (double)v.x * inv;
      v.y = (double)v.y * inv;
      x[i] = v;
    }
    if (tid == 0) acc[2] += log(m);
  }
}

int launch_rescale_log(void* x, int64_t count, double* acc, int dtype, hipStream_t s) {
  if (dtype == TENET_C128)
    hipLaunchKernelGGL(rescale_log_kernel<double2>, dim3(1), dim3(1024), 0, s, (double2*)x, count, acc);
  else
    hipLaunchKernelGGL(rescale_log_kernel<float2>, dim3(1), dim3(1024), 0, s, (float2*)x, count, acc);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

template <typename S2>
__global__ void write_scalar_kernel(const S2* env, double* out) {
  out[0] = (double)env[0].x;
  out[1] = (double)env[0].y;
}
int launch_write_scalar_from(const void* env, double* out, int dtype, hipStream_t s) {
  if (dtype == TENET_C128)
    hipLaunchKernelGGL(write_scalar_kernel<double2>, dim3(1), dim3(1), 0, s, (const double2*)env, out);
  else
    hipLaunchKernelGGL(write_scalar_kernel<float2>, dim3(1), dim3(1), 0, s, (const float2*)env, out);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

// <b|a> for small bonds in one workgroup (overlap.jl:36-50): the environment
// E (chi_a x chi_b) stays in LDS while the sites stream through, so a chain
// costs one launch instead of two contractions and a rescale per site.
// Per site i (E = [1] before site 0):
//   T(lb, s, ra)  = sum_la E(la, lb) A(la, s, ra)           (lb x d*ra, K = la)
//   E'(ra, rb)    = sum_{lb, s} T(lb, s, ra) conj(B(lb, s, rb))  (K = lb*d)
// then E' is divided by its largest |re| / |im| and the log of that scale is
// added to out[2], as the multi-launch sweep does.  FP64 accumulation; each
// of the 512 threads holds a 2 x 2 block of T and a 2 x 1 block of E' (two
// waves per SIMD); A, B and T sit in LDS transposed so that the lanes of a
// wave read consecutive entries, and the next site's loads are in flight
// while the current one is contracted.
constexpr int OS_THREADS = 512;  // 1024 threads with 1 x 2 / 1 x 1 blocks measured no faster
constexpr int OS_TR = 2;         // T block: OS_TR rows x 2 columns per thread
constexpr int OS_ER = 2;         // E' block: OS_ER rows x 1 column per thread
constexpr int OS_CHI = 32;     // every bond <= 32
constexpr int OS_SITE = 2048;  // d * chi^2 <= 2048 (A, B and T staged in LDS)
constexpr int OS_PER = OS_SITE / OS_THREADS;

template <typename S2>
__device__ __forceinline__ void os_load(const OverlapTab& tab, int i, int tid, S2 (&va)[OS_PER], S2 (&vb)[OS_PER]) {
  const int d = tab.d;
  const int na = (int)(tab.chi_a[i] * d * tab.chi_a[i + 1]), nb = (int)(tab.chi_b[i] * d * tab.chi_b[i + 1]);
  const S2* A = reinterpret_cast<const S2*>(tab.a[i]);
  const S2* B = reinterpret_cast<const S2*>(tab.b[i]);
#pragma unroll
  for (int u = 0; u < OS_PER; ++u) {
    const int e = tid + u * OS_THREADS;
    if (e < na) va[u] = A[e];
    if (e < nb) vb[u] = B[e];
  }
}

template <typename S2>
__global__ __launch_bounds__(OS_THREADS) void overlap_small_kernel(const OverlapTab tab, double* __restrict__ out) {
  __shared__ double Er[OS_CHI][OS_CHI + 1], Ei[OS_CHI][OS_CHI + 1];  // E(la, lb)
  __shared__ double Ar[OS_SITE], Ai[OS_SITE], Br[OS_SITE], Bi[OS_SITE], Tr[OS_SITE], Ti[OS_SITE];
  __shared__ double red[OS_THREADS / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int n = tab.n, d = tab.d;
  const int64_t* chi_a = tab.chi_a;
  const int64_t* chi_b = tab.chi_b;
  if (tid == 0) {
    Er[0][0] = 1.0;
    Ei[0][0] = 0.0;
  }
  double logacc = 0.0;
  S2 va[OS_PER], vb[OS_PER];
  os_load(tab, 0, tid, va, vb);
  __syncthreads();
  for (int i = 0; i < n; ++i) {
    const int la = (int)chi_a[i], lb = (int)chi_b[i], ra = (int)chi_a[i + 1], rb = (int)chi_b[i + 1];
    const int na = la * d * ra, nb = lb * d * rb;
    const int N = d * ra, K = lb * d;
    // stored transposed so that the index running across lanes in the
    // products below is the contiguous one (no LDS bank conflicts):
    //   A'[c + N a] = A(a, c),  c = s + d ra     B'[rb + rb_n k] = B(k, rb),  k = lb + lb_n s
#pragma unroll
    for (int u = 0; u < OS_PER; ++u) {
      const int e = tid + u * OS_THREADS;
      if (e < na) {
        const int ia = e % la, ic = e / la;
        Ar[ic + N * ia] = (double)va[u].x;
        Ai[ic + N * ia] = (double)va[u].y;
      }
      if (e < nb) {
        const int ik = e % K, ib = e / K;
        Br[ib + rb * ik] = (double)vb[u].x;
        Bi[ib + rb * ik] = (double)vb[u].y;
      }
    }
    __syncthreads();
    if (i + 1 < n) os_load(tab, i + 1, tid, va, vb);  // in flight during the products
    // T = E^T A: rows lb, columns c = s + d * ra; stored as T'[ra + ra_n (lb + lb_n s)]
    {
      const int M = lb, mb = (M + OS_TR - 1) / OS_TR, nbk = (N + 1) / 2;
      for (int blk = tid; blk < mb * nbk; blk += OS_THREADS) {
        const int r0 = OS_TR * (blk % mb), c0 = 2 * (blk / mb);
        double tr[OS_TR][2], ti[OS_TR][2];
#pragma unroll
        for (int x = 0; x < OS_TR; ++x)
#pragma unroll
          for (int y = 0; y < 2; ++y) tr[x][y] = ti[x][y] = 0.0;
        for (int a = 0; a < la; ++a) {
          double er[OS_TR], ei[OS_TR], xr[2], xi[2];
#pragma unroll
          for (int x = 0; x < OS_TR; ++x) {
            const int r = min(r0 + x, M - 1);
            er[x] = Er[a][r];
            ei[x] = Ei[a][r];
          }
#pragma unroll
          for (int y = 0; y < 2; ++y) {
            const int c = min(c0 + y, N - 1);
            xr[y] = Ar[c + N * a];
            xi[y] = Ai[c + N * a];
          }
#pragma unroll
          for (int x = 0; x < OS_TR; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y) {
              tr[x][y] = fma(er[x], xr[y], fma(-ei[x], xi[y], tr[x][y]));
              ti[x][y] = fma(er[x], xi[y], fma(ei[x], xr[y], ti[x][y]));
            }
        }
#pragma unroll
        for (int x = 0; x < OS_TR; ++x)
#pragma unroll
          for (int y = 0; y < 2; ++y)
            if (r0 + x < M && c0 + y < N) {
              const int c = c0 + y, sidx = c % d, rai = c / d;
              const int t = rai + ra * ((r0 + x) + lb * sidx);
              Tr[t] = tr[x][y];
              Ti[t] = ti[x][y];
            }
      }
    }
    __syncthreads();
    // E'(ra, rb) = sum_k T'[ra + ra_n k] conj(B'[rb + rb_n k]), k = lb + lb_n s
    const int mb = (ra + OS_ER - 1) / OS_ER;
    double outr[OS_ER], outi[OS_ER];
#pragma unroll
    for (int x = 0; x < OS_ER; ++x) outr[x] = outi[x] = 0.0;
    const bool has = tid < mb * rb;  // mb * rb <= OS_THREADS
    const int a0 = OS_ER * (tid % max(mb, 1)), b0 = tid / max(mb, 1);
    double mx = 0.0;
    if (has) {
      for (int k = 0; k < K; ++k) {
        const double brr = Br[b0 + rb * k], bii = Bi[b0 + rb * k];
#pragma unroll
        for (int x = 0; x < OS_ER; ++x) {  // T * conj(B)
          const int ai = min(a0 + x, ra - 1);
          const double trr = Tr[ai + ra * k], tii = Ti[ai + ra * k];
          outr[x] = fma(trr, brr, fma(tii, bii, outr[x]));
          outi[x] = fma(tii, brr, fma(-trr, bii, outi[x]));
        }
      }
#pragma unroll
      for (int x = 0; x < OS_ER; ++x)
        if (a0 + x < ra) mx = fmax(mx, fmax(fabs(outr[x]), fabs(outi[x])));
    }
    mx = wave_max(mx);
    if (lane == 0) red[wave] = mx;
    __syncthreads();  // also: every read of E, A, T above is done
    double m = red[0];
#pragma unroll
    for (int w = 1; w < OS_THREADS / 64; ++w) m = fmax(m, red[w]);
    const double inv = m > 0.0 ? 1.0 / m : 1.0;
    if (has) {
#pragma unroll
      for (int x = 0; x < OS_ER; ++x)
        if (a0 + x < ra) {
          Er[a0 + x][b0] = outr[x] * inv;
          Ei[a0 + x][b0] = outi[x] * inv;
        }
    }
    if (m > 0.0) logacc += log(m);
    __syncthreads();
  }
  if (tid == 0) {
    out[0] = Er[0][0];
    out[1] = Ei[0][0];
    out[2] = logacc;
  }
}

bool overlap_small_fits(int64_t n, int64_t d, const int64_t* chi_a, const int64_t* chi_b) {
  if (n < 1 || n > OS_NMAX || d < 1) return false;
  int64_t mx = 1;
  for (int64_t i = 0; i <= n; ++i) mx = std::max(mx, std::max(chi_a[i], chi_b[i]));
  return mx <= OS_CHI && d * mx * mx <= OS_SITE;
}

int launch_overlap_small(const OverlapTab& tab, double* out, int dtype, hipStream_t s) {
  if (dtype == TENET_C128)
    hipLaunchKernelGGL(overlap_small_kernel<double2>, dim3(1), dim3(OS_THREADS), 0, s, tab, out);
  else
    hipLaunchKernelGGL(overlap_small_kernel<float2>, dim3(1), dim3(OS_THREADS), 0, s, tab, out);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

// out[0:2] = trace(env) for a chi x chi column-major env
template <typename S2>
__global__ __launch_bounds__(256) void trace_kernel(const S2* env, int64_t chi, double* out) {
  __shared__ double rr[4], ri[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double sr = 0.0, si = 0.0;
  for (int64_t i = tid; i < chi; i += 256) {
    const S2 v = env[i * chi + i];
    sr += (double)v.x;
    si += (double)v.y;
  }
  sr = wave_add(sr);
  si = wave_add(si);
  if (lane == 0) {
    rr[wave] = sr;
    ri[wave] = si;
  }
  __syncthreads();
  if (tid == 0) {
    out[0] = rr[0] + rr[1] + rr[2] + rr[3];
    out[1] = ri[0] + ri[1] + ri[2] + ri[3];
  }
}
int launch_trace_diag(const void* env, int64_t chi, double* out, int dtype, hipStream_t s) {
  if (dtype == TENET_C128)
    hipLaunchKernelGGL(trace_kernel<double2>, dim3(1), dim3(256), 0, s, (const double2*)env, chi, out);
  else
    hipLaunchKernelGGL(trace_kernel<float2>, dim3(1), dim3(256), 0, s, (const float2*)env, chi, out);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

template <typename SI, typename SO>
__global__ void convert_kernel(const SI* __restrict__ src, SO* __restrict__ dst, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const SI v = src[i];
    SO o;
    o.x = v.x;
    o.y = v.y;
    dst[i] = o;
  }
}
int launch_convert(const void* src, int src_dtype, void* dst, int dst_dtype, int64_t count, hipStream_t s) {
  const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>((count + 255) / 256, 4096));
  if (src_dtype == TENET_C128 && dst_dtype == TENET_C64)
    hipLaunchKernelGGL((convert_kernel<double2, float2>), dim3(g), dim3(256), 0, s, (const double2*)src, (float2*)dst,
                       count);
  else if (src_dtype == TENET_C64 && dst_dtype == TENET_C128)
    hipLaunchKernelGGL((convert_kernel<float2, double2>), dim3(g), dim3(256), 0, s, (const float2*)src, (double2*)dst,
                       count);
  else if (src_dtype == TENET_C128)
    hipLaunchKernelGGL((convert_kernel<double2, double2>), dim3(g), dim3(256), 0, s, (const double2*)src,
                       (double2*)dst, count);
  else
    hipLaunchKernelGGL((convert_kernel<float2, float2>), dim3(g), dim3(256), 0, s, (const float2*)src, (float2*)dst,
                       count);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

// ---- K4: batched amplitude sweep ------------------------------------------------
// One workgroup per request.  The running row vector v (chi <= max_chi) lives
// in LDS as FP64; per site v'[r] = sum_l v[l] A[l, x, r] with A stored
// (chi_l, d, chi_r) column-major, so the slice A[:, x, :] is a chi_l x chi_r
// matrix of leading dimension chi_l*d whose columns are contiguous.  Each wave
// takes 8 columns at a time: its lanes stride over l (coalesced loads, eight
// independent streams in flight) and the 16 partial sums are combined by one
// butterfly transpose-reduction (17 shuffles, not 16 full reductions).  The
// next vector's maximum comes out of the same pass; v is rescaled lazily by
// its inverse when it is read at the next site, and the log of every scale
// is accumulated in FP64 (unnormalised states overflow otherwise).  One
// barrier per site.
constexpr int AMP_THREADS = 256;
constexpr int AMP_MAX_CHI = 2048;
constexpr int AMP_COLS = 8;

template <int H, int BIT>
__device__ __forceinline__ void amp_halve(double (&v)[2 * AMP_COLS], int lane) {
  const bool hi = (lane & BIT) != 0;
#pragma unroll
  for (int i = 0; i < H; ++i) {
    const double keep = hi ? v[H + i] : v[i];
    const double send = hi ? v[i] : v[H + i];
    v[i] = keep + __shfl_xor(send, BIT, 64);
  }
}

template <typename S2>
__global__ __launch_bounds__(AMP_THREADS) void amplitude_kernel(const int64_t* __restrict__ n_dev,
                                                                 const uint64_t* __restrict__ site_ptrs,
                                                                 const uint64_t* __restrict__ chis,
                                                                 const uint64_t* __restrict__ bits, int64_t d,
                                                                 double* __restrict__ out) {
  __shared__ double vr[2][AMP_MAX_CHI], vi[2][AMP_MAX_CHI];
  __shared__ double wmax[2][AMP_THREADS / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr int NWAVE = AMP_THREADS / 64;
  const int64_t req = blockIdx.x;
  const int64_t n = n_dev[req];
  const uint64_t* const sites = reinterpret_cast<const uint64_t*>(site_ptrs[req]);
  const int64_t* const chi = reinterpret_cast<const int64_t*>(chis[req]);
  const uint8_t* const x = reinterpret_cast<const uint8_t*>(bits[req]);
  if (tid == 0) {
    vr[0][0] = 1.0;
    vi[0][0] = 0.0;
  }
  double logscale = 0.0, scale = 1.0;
  int cur = 0;
  __syncthreads();
  for (int64_t i = 0; i < n; ++i) {
    const int64_t cl = chi[i], cr = chi[i + 1];
    if (cl < 1 || cr < 1 || cl > AMP_MAX_CHI || cr > AMP_MAX_CHI || (int64_t)x[i] >= d) {
      // uniform across the workgroup: report NaN instead of reading out of bounds
      if (tid == 0) {
        out[3 * req + 0] = __builtin_nan("");
        out[3 * req + 1] = __builtin_nan("");
        out[3 * req + 2] = 0.0;
      }
      return;
    }
    const S2* A = reinterpret_cast<const S2*>(sites[i]) + (int64_t)x[i] * cl;
    const int64_t ld = cl * d;
    double lmax = 0.0;
    for (int64_t cb = (int64_t)wave * AMP_COLS; cb < cr; cb += (int64_t)NWAVE * AMP_COLS) {
      double acc[2 * AMP_COLS];
#pragma unroll
      for (int j = 0; j < 2 * AMP_COLS; ++j) acc[j] = 0.0;
      for (int64_t l = lane; l < cl; l += 64) {
        const double pr = vr[cur][l] * scale, pi = vi[cur][l] * scale;
#pragma unroll
        for (int j = 0; j < AMP_COLS; ++j) {
          if (cb + j < cr) {
            const S2 a = A[l + (cb + j) * ld];
            const double ar = (double)a.x, ai = (double)a.y;
            acc[2 * j] += pr * ar - pi * ai;
            acc[2 * j + 1] += pr * ai + pi * ar;
          }
        }
      }
      // lane l ends with the wave total of value l >> 2 (four lanes each)
      amp_halve<8, 32>(acc, lane);
      amp_halve<4, 16>(acc, lane);
      amp_halve<2, 8>(acc, lane);
      amp_halve<1, 4>(acc, lane);
      acc[0] += __shfl_xor(acc[0], 2, 64);
      acc[0] += __shfl_xor(acc[0], 1, 64);
      const int k = lane >> 2;
      const int64_t col = cb + (k >> 1);
      if ((lane & 3) == 0 && col < cr) {
        if (k & 1)
          vi[cur ^ 1][col] = acc[0];
        else
          vr[cur ^ 1][col] = acc[0];
        lmax = fmax(lmax, fabs(acc[0]));
      }
    }
    lmax = wave_max(lmax);
    if (lane == 0) wmax[cur][wave] = lmax;
    __syncthreads();
    double m = 0.0;
#pragma unroll
    for (int w = 0; w < NWAVE; ++w) m = fmax(m, wmax[cur][w]);
    cur ^= 1;
    if (m > 0.0) {
      scale = 1.0 / m;
      logscale += log(m);
    } else {
      scale = 1.0;
    }
  }
  if (tid == 0) {
    out[3 * req + 0] = vr[cur][0] * scale;
    out[3 * req + 1] = vi[cur][0] * scale;
    out[3 * req + 2] = logscale;
  }
}

// ---- site-parallel amplitudes ------------------------------------------------
// One launch per site position: workgroup (column block, request) computes
// AS_COLS entries of v' = (v / m) A[:, x_i, :] for its request, where v is the
// request's previous vector (read whole from the workspace into LDS) and m
// its largest |component| (every workgroup of the request computes the same
// m, so the rescale is deterministic); the workgroup of column block 0 adds
// log(m) to the request's log-scale.  A request's vectors alternate between
// two workspace rows; at its last site (chi_r = 1) block 0 writes
// (re, im, log-scale) to out.  The chain is sequential over sites, so a
// request's sites are spread over launches instead of one workgroup, and a
// 512 x 512 slice is read by 32 workgroups at once.
constexpr int AS_THREADS = 256;
constexpr int AS_COLS = 16;  // columns per workgroup: 4 per wave

template <typename S2>
__global__ __launch_bounds__(AS_THREADS) void amp_site_kernel(AmpSiteArgs a, int64_t site) {
  __shared__ double vr[AMP_MAX_CHI], vi[AMP_MAX_CHI];
  __shared__ double wm[AS_THREADS / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t req = blockIdx.y;
  const int64_t n = a.n[req];
  if (n > a.n_max) {
    // more sites than the launches cover: report NaN once
    if (site == 0 && blockIdx.x == 0 && threadIdx.x == 0) {
      a.out[3 * req + 0] = __builtin_nan("");
      a.out[3 * req + 1] = __builtin_nan("");
      a.out[3 * req + 2] = 0.0;
    }
    return;
  }
  if (site >= n) return;
  const int64_t* chi = reinterpret_cast<const int64_t*>(a.chis[req]);
  const uint8_t* x = reinterpret_cast<const uint8_t*>(a.bits[req]);
  const int64_t cl = chi[site], cr = chi[site + 1];
  const int64_t cb0 = (int64_t)blockIdx.x * AS_COLS;
  if (cb0 >= cr && !(cr < 1 && blockIdx.x == 0)) return;
  double* out = a.out + 3 * req;
  if (cl < 1 || cr < 1 || cl > a.chi_max || cr > a.chi_max || (int64_t)x[site] >= a.d || (site == 0 && cl != 1) ||
      (site == n - 1 && cr != 1)) {
    // uniform across the request's workgroups: report NaN instead of reading
    // out of bounds, and mark the request so that its later sites skip it
    if (blockIdx.x == 0 && tid == 0) {
      out[0] = __builtin_nan("");
      out[1] = __builtin_nan("");
      out[2] = 0.0;
      a.logs[req] = __builtin_nan("");
    }
    return;
  }
  if (site > 0 && __builtin_isnan(a.logs[req])) return;  // marked at an earlier site
  double* vin = a.vbuf + ((size_t)req * 2 + ((site + 1) & 1)) * 2 * a.chi_max;
  double* vout = a.vbuf + ((size_t)req * 2 + (site & 1)) * 2 * a.chi_max;
  // previous vector into LDS and its largest component
  double m = 1.0;
  if (site == 0) {
    if (tid == 0) {
      vr[0] = 1.0;
      vi[0] = 0.0;
    }
  } else {
    double mx = 0.0;
    for (int64_t l = tid; l < cl; l += AS_THREADS) {
      const double re = vin[2 * l], im = vin[2 * l + 1];
      vr[l] = re;
      vi[l] = im;
      mx = fmax(mx, fmax(fabs(re), fabs(im)));
    }
    mx = wave_max(mx);
    if (lane == 0) wm[wave] = mx;
    __syncthreads();
    m = fmax(fmax(wm[0], wm[1]), fmax(wm[2], wm[3]));
    if (!(m > 0.0)) m = 1.0;
  }
  __syncthreads();
  const double sc = 1.0 / m;
  if (blockIdx.x == 0 && tid == 0) a.logs[req] = (site == 0 ? 0.0 : a.logs[req] + log(m));
  const S2* A = reinterpret_cast<const S2*>(reinterpret_cast<const uint64_t*>(a.sites[req])[site]) +
                (int64_t)x[site] * cl;
  const int64_t ld = cl * a.d;
  const int64_t c0 = cb0 + 4 * wave;
  double acc[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) acc[j] = 0.0;
#pragma unroll 2
  for (int64_t l = lane; l < cl; l += 64) {
    const double pr = vr[l] * sc, pi = vi[l] * sc;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (c0 + j < cr) {
        const S2 v = A[l + (c0 + j) * ld];
        const double ar = (double)v.x, ai = (double)v.y;
        acc[2 * j] += pr * ar - pi * ai;
        acc[2 * j + 1] += pr * ai + pi * ar;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc[j] += __shfl_xor(acc[j], off, 64);
  }
  if (lane < 4 && c0 + lane < cr) {
    double re = acc[0], im = acc[1];
#pragma unroll
    for (int j = 1; j < 4; ++j)
      if (lane == j) {
        re = acc[2 * j];
        im = acc[2 * j + 1];
      }
    vout[2 * (c0 + lane)] = re;
    vout[2 * (c0 + lane) + 1] = im;
    if (site == n - 1 && c0 + lane == 0) {
      // the last site has chi_r = 1: its single entry is the amplitude
      out[0] = re;
      out[1] = im;
      out[2] = a.logs[req];
    }
  }
}

int launch_amp_site(const AmpSiteArgs& a, int64_t nreq, int64_t site, int dtype, hipStream_t s, hipEvent_t e0,
                    hipEvent_t e1) {
  if (nreq <= 0) return TENET_OK;
  if (a.chi_max > AMP_MAX_CHI) {
    set_error("amplitudes: bond dimension above 2048");
    return TENET_ERR_UNSUPPORTED;
  }
  const dim3 grid((unsigned)((a.chi_max + AS_COLS - 1) / AS_COLS), (unsigned)nreq);
  if (dtype == TENET_C128) {
    if (e0)
      hipExtLaunchKernelGGL(amp_site_kernel<double2>, grid, dim3(AS_THREADS), 0, s, e0, e1, 0, a, site);
    else
      hipLaunchKernelGGL(amp_site_kernel<double2>, grid, dim3(AS_THREADS), 0, s, a, site);
  } else {
    if (e0)
      hipExtLaunchKernelGGL(amp_site_kernel<float2>, grid, dim3(AS_THREADS), 0, s, e0, e1, 0, a, site);
    else
      hipLaunchKernelGGL(amp_site_kernel<float2>, grid, dim3(AS_THREADS), 0, s, a, site);
  }
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

int launch_amplitudes(int64_t nreq, const int64_t* n_dev, const uint64_t* site_ptrs, const uint64_t* chis,
                      const uint64_t* bits, int64_t d, double* out, int dtype, int max_chi, hipStream_t s) {
  if (nreq <= 0) return TENET_OK;
  if (max_chi > AMP_MAX_CHI) {
    set_error("amplitudes: bond dimension above 2048");
    return TENET_ERR_UNSUPPORTED;
  }
  if (dtype == TENET_C128)
    hipLaunchKernelGGL(amplitude_kernel<double2>, dim3((unsigned)nreq), dim3(AMP_THREADS), 0, s, n_dev, site_ptrs,
                       chis, bits, d, out);
  else
    hipLaunchKernelGGL(amplitude_kernel<float2>, dim3((unsigned)nreq), dim3(AMP_THREADS), 0, s, n_dev, site_ptrs,
                       chis, bits, d, out);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

// dst (m x n, ld m, complex FP64) <- src (m x n, ld lda, dtype), or, with
// adj, <- src^H where src is n x m with leading dimension lda.
template <typename S2>
__global__ void pack_c128_kernel(const S2* __restrict__ src, int64_t m, int64_t n, int64_t lda, int adj,
                                 double2* __restrict__ dst) {
  const size_t total = (size_t)m * n;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const int64_t r = (int64_t)(e % m), c = (int64_t)(e / m);
    double2 o;
    if (!adj) {
      const S2 v = src[(size_t)c * lda + r];
      o.x = v.x;
      o.y = v.y;
    } else {
      const S2 v = src[(size_t)r * lda + c];
      o.x = v.x;
      o.y = -(double)v.y;
    }
    dst[e] = o;
  }
}
int launch_pack_c128(const void* src, int dtype, int64_t m, int64_t n, int64_t lda, int adj, void* dst,
                     hipStream_t s) {
  const size_t total = (size_t)m * n;
  const unsigned g = (unsigned)std::max<size_t>(1, std::min<size_t>((total + 255) / 256, 4096));
  if (dtype == TENET_C128)
    hipLaunchKernelGGL(pack_c128_kernel<double2>, dim3(g), dim3(256), 0, s, (const double2*)src, m, n, lda, adj,
                       (double2*)dst);
  else
    hipLaunchKernelGGL(pack_c128_kernel<float2>, dim3(g), dim3(256), 0, s, (const float2*)src, m, n, lda, adj,
                       (double2*)dst);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

// out (chil, 2, chir) = P applied to the physical index of A (chil, 2, chir)
// Pauli matrices as in ext/TenetPythonCallExt/pytket.jl:23-31.
template <typename S2>
__global__ void pauli_kernel(const S2* __restrict__ A, int64_t chil, int64_t chir, int which, S2* __restrict__ out) {
  const size_t total = (size_t)chil * 2 * chir;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const int64_t l = (int64_t)(e % chil);
    const int64_t o = (int64_t)((e / chil) % 2);
    const int64_t r = (int64_t)(e / (2 * chil));
    const S2 a0 = A[l + chil * (0 + 2 * r)];
    const S2 a1 = A[l + chil * (1 + 2 * r)];
    S2 v;
    if (which == 'X') {
      v = o == 0 ? a1 : a0;
    } else if (which == 'Y') {  // [[0, -i], [i, 0]]
      if (o == 0) {
        v.x = a1.y;
        v.y = -a1.x;
      } else {
        v.x = -a0.y;
        v.y = a0.x;
      }
    } else if (which == 'Z') {
      v = o == 0 ? a0 : a1;
      if (o == 1) {
        v.x = -v.x;
        v.y = -v.y;
      }
    } else {
      v = o == 0 ? a0 : a1;
    }
    out[e] = v;
  }
}
int launch_apply_pauli(const void* A, int64_t chil, int64_t chir, char which, void* out, int dtype, hipStream_t s) {
  const size_t total = (size_t)chil * 2 * chir;
  const unsigned g = (unsigned)std::max<size_t>(1, std::min<size_t>((total + 255) / 256, 4096));
  if (dtype == TENET_C128)
    hipLaunchKernelGGL(pauli_kernel<double2>, dim3(g), dim3(256), 0, s, (const double2*)A, chil, chir, (int)which,
                       (double2*)out);
  else
    hipLaunchKernelGGL(pauli_kernel<float2>, dim3(g), dim3(256), 0, s, (const float2*)A, chil, chir, (int)which,
                       (float2*)out);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

template <typename S2>
__global__ void identity_kernel(S2* __restrict__ out, int64_t chi) {
  const size_t total = (size_t)chi * chi;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    S2 v;
    v.x = ((int64_t)(e % chi) == (int64_t)(e / chi)) ? 1 : 0;
    v.y = 0;
    out[e] = v;
  }
}
int launch_set_identity(void* out, int64_t chi, int dtype, hipStream_t s) {
  const size_t total = (size_t)chi * chi;
  const unsigned g = (unsigned)std::max<size_t>(1, std::min<size_t>((total + 255) / 256, 4096));
  if (dtype == TENET_C128)
    hipLaunchKernelGGL(identity_kernel<double2>, dim3(g), dim3(256), 0, s, (double2*)out, chi);
  else
    hipLaunchKernelGGL(identity_kernel<float2>, dim3(g), dim3(256), 0, s, (float2*)out, chi);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

// Counter-based standard complex Gaussians for synthetic workloads:
// element i gets u1 = U(h(2i)), u2 = U(h(2i+1)), h(k) = splitmix64 finaliser of
// (k * 0x9E3779B97F4A7C15 + seed * 0xD1B54A32D192ED03), U(x) = (x >> 11) 2^-53;
// z = sqrt(-2 ln(1 - u1)) (cos 2 pi u2 + i sin 2 pi u2) / sqrt(2)  (Box-Muller).
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
template <typename S2>
__global__ void gauss_kernel(S2* __restrict__ out, int64_t count, uint64_t seed, uint64_t offset) {
  const uint64_t sm = seed * 0xD1B54A32D192ED03ull;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t k = 2ull * ((uint64_t)i + offset);
    const double u1 = (double)(mix64(k * 0x9E3779B97F4A7C15ull + sm) >> 11) * 0x1p-53;
    const double u2 = (double)(mix64((k + 1) * 0x9E3779B97F4A7C15ull + sm) >> 11) * 0x1p-53;
    const double r = sqrt(-2.0 * log(1.0 - u1)) * 0.70710678118654752440;
    double sn, cs;
    sincospi(2.0 * u2, &sn, &cs);
    S2 v;
    v.x = r * cs;
    v.y = r * sn;
    out[i] = v;
  }
}
int launch_gaussian(void* out, int dtype, int64_t count, uint64_t seed, uint64_t offset, hipStream_t s) {
  if (count <= 0) return TENET_OK;
  const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>((count + 255) / 256, 16384));
  if (dtype == TENET_C128)
    hipLaunchKernelGGL(gauss_kernel<double2>, dim3(g), dim3(256), 0, s, (double2*)out, count, seed, offset);
  else
    hipLaunchKernelGGL(gauss_kernel<float2>, dim3(g), dim3(256), 0, s, (float2*)out, count, seed, offset);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

// One sampling step at one site (Tenet's sample, src/Operations/sample.jl:20-72).
// T holds every sample's projected site tensor, T[x, r, s] = sum_l V[l, s]
// A[l, x, r] (column-major d x chir x nsamp).  One wave per sample:
// p_x = sum_r |T[x, r, s]|^2 in FP64 (all lanes end with the same totals: an
// xor butterfly adds the same pairs in every lane), the sample's uniform u is
// element s * n + site of the generator above (u1 of that element), x* is the
// first x whose running sum of p exceeds u * sum_x p_x (the last x with
// p_x > 0 when rounding leaves none), and V'[r, s] = T[x*, r, s] / sqrt(p_x*).
constexpr int SAMPLE_MAX_D = 8;
template <typename S2>
__global__ __launch_bounds__(256) void sample_select_kernel(const S2* __restrict__ T, int d, int chir, int64_t nsamp,
                                                            int n, int site, uint64_t seed, int32_t* __restrict__ bits,
                                                            S2* __restrict__ Vn) {
  const int lane = threadIdx.x & 63;
  const int64_t s = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (s >= nsamp) return;  // whole waves; the kernel has no barrier
  const S2* t = T + s * (int64_t)d * chir;
  double p[SAMPLE_MAX_D];
#pragma unroll
  for (int x = 0; x < SAMPLE_MAX_D; ++x) p[x] = 0.0;
  for (int r = lane; r < chir; r += 64) {
#pragma unroll
    for (int x = 0; x < SAMPLE_MAX_D; ++x) {
      if (x < d) {
        const S2 v = t[x + (int64_t)d * r];
        p[x] += (double)v.x * v.x + (double)v.y * v.y;
      }
    }
  }
#pragma unroll
  for (int x = 0; x < SAMPLE_MAX_D; ++x)
    for (int off = 32; off > 0; off >>= 1) p[x] += __shfl_xor(p[x], off, 64);
  double total = 0.0;
  for (int x = 0; x < d; ++x) total += p[x];
  const uint64_t k = 2ull * ((uint64_t)s * (uint64_t)n + (uint64_t)site);
  const double u = (double)(mix64(k * 0x9E3779B97F4A7C15ull + seed * 0xD1B54A32D192ED03ull) >> 11) * 0x1p-53;
  const double target = u * total;
  int pick = -1, last = 0;
  double cum = 0.0;
  for (int x = 0; x < d; ++x) {
    if (p[x] > 0.0) last = x;
    cum += p[x];
    if (pick < 0 && p[x] > 0.0 && cum > target) pick = x;
  }
  if (pick < 0) pick = last;
  if (lane == 0) bits[s * n + site] = pick;
  if (Vn) {
    const double sc = p[pick] > 0.0 ? 1.0 / sqrt(p[pick]) : 0.0;
    for (int r = lane; r < chir; r += 64) {
      const S2 v = t[pick + (int64_t)d * r];
      S2 o;
      o.x = v.x * sc;
      o.y = v.y * sc;
      Vn[r + (int64_t)chir * s] = o;
    }
  }
}
int launch_sample_select(const void* T, int dtype, int d, int chir, int64_t nsamp, int n, int site, uint64_t seed,
                         int32_t* bits, void* Vn, hipStream_t s) {
  if (nsamp <= 0) return TENET_OK;
  const unsigned g = (unsigned)((nsamp + 3) / 4);
  if (dtype == TENET_C128)
    hipLaunchKernelGGL(sample_select_kernel<double2>, dim3(g), dim3(256), 0, s, (const double2*)T, d, chir, nsamp, n,
                       site, seed, bits, (double2*)Vn);
  else
    hipLaunchKernelGGL(sample_select_kernel<float2>, dim3(g), dim3(256), 0, s, (const float2*)T, d, chir, nsamp, n,
                       site, seed, bits, (float2*)Vn);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

// x *= alpha (complex) for count elements
template <typename S2>
__global__ void scale_kernel(S2* __restrict__ x, int64_t count, double ar, double ai) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
    const S2 v = x[i];
    S2 o;
    o.x = ar * (double)v.x - ai * (double)v.y;
    o.y = ar * (double)v.y + ai * (double)v.x;
    x[i] = o;
  }
}
int launch_scale(void* x, int dtype, int64_t count, double ar, double ai, hipStream_t s) {
  if (count <= 0) return TENET_OK;
  const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>((count + 255) / 256, 16384));
  if (dtype == TENET_C128)
    hipLaunchKernelGGL(scale_kernel<double2>, dim3(g), dim3(256), 0, s, (double2*)x, count, ar, ai);
  else
    hipLaunchKernelGGL(scale_kernel<float2>, dim3(g), dim3(256), 0, s, (float2*)x, count, ar, ai);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

// sum |x|^2 in FP64: per-block partials then one block; fixed order
template <typename S2>
__global__ __launch_bounds__(256) void nrm2_partial_kernel(const S2* __restrict__ x, int64_t count, double* part) {
  __shared__ double red[4];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
    const S2 v = x[i];
    s += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
  }
  s = wave_add(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}
__global__ __launch_bounds__(256) void nrm2_final_kernel(const double* part, int n, double* out) {
  __shared__ double red[4];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) s += part[i];
  s = wave_add(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[0] = red[0] + red[1] + red[2] + red[3];
}
int launch_nrm2(const void* x, int dtype, int64_t count, double* part, double* out, hipStream_t s) {
  constexpr int NB = 1024;
  if (dtype == TENET_C128)
    hipLaunchKernelGGL(nrm2_partial_kernel<double2>, dim3(NB), dim3(256), 0, s, (const double2*)x, count, part);
  else
    hipLaunchKernelGGL(nrm2_partial_kernel<float2>, dim3(NB), dim3(256), 0, s, (const float2*)x, count, part);
  TENET_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(nrm2_final_kernel, dim3(1), dim3(256), 0, s, part, NB, out);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

// sq[j] = sqrt(S[j]), isq[j] = 1 / sqrt(S[j]) (0 where S[j] = 0), as complex
template <typename S2>
__global__ void sigma_scales_kernel(const double* __restrict__ S, int64_t k, S2* __restrict__ sq, S2* __restrict__ isq) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < k; j += (int64_t)gridDim.x * blockDim.x) {
    const double r = sqrt(fmax(S[j], 0.0));
    S2 a, b;
    a.x = r;
    a.y = 0.0;
    b.x = r > 0.0 ? 1.0 / r : 0.0;
    b.y = 0.0;
    sq[j] = a;
    isq[j] = b;
  }
}
int launch_sigma_scales(const double* S, int64_t k, void* sq, void* isq, int dtype, hipStream_t s) {
  if (k <= 0) return TENET_OK;
  const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>((k + 255) / 256, 64));
  if (dtype == TENET_C128)
    hipLaunchKernelGGL(sigma_scales_kernel<double2>, dim3(g), dim3(256), 0, s, S, k, (double2*)sq, (double2*)isq);
  else
    hipLaunchKernelGGL(sigma_scales_kernel<float2>, dim3(g), dim3(256), 0, s, S, k, (float2*)sq, (float2*)isq);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

// Two-site gate application (the gate contraction of a simple update,
// batched over the disjoint bonds of a brickwork layer).  One thread per
// (l, r) of a problem reads the d^2 entries theta[l, :, :, r] once and
// writes all d^2 outputs, so theta leaves memory once; for fixed (x, y)
// consecutive threads touch consecutive l (coalesced).  The gate is held in
// registers.  R: the accumulation type (double for complex128, float for
// complex64, as the library's other complex64 contractions).
template <typename S2, typename R, int D>
__global__ void gate_apply_kernel(GateBatch B) {
  const GateProb& P = B.p[blockIdx.y];
  const int64_t n = P.chil * P.chir;
  const S2* __restrict__ th = static_cast<const S2*>(P.theta);
  const S2* __restrict__ G = static_cast<const S2*>(P.gate);
  const S2* lam = static_cast<const S2*>(P.lam);
  S2* __restrict__ to = static_cast<S2*>(P.theta_out);
  S2* xo = static_cast<S2*>(P.x_out);
  R gr[D * D * D * D], gi[D * D * D * D];
#pragma unroll
  for (int e = 0; e < D * D * D * D; ++e) {
    const S2 g = G[e];
    gr[e] = (R)g.x;
    gi[e] = (R)g.y;
  }
  const int64_t ld = P.chil;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t l = e % ld, r = e / ld;
    const int64_t base = l + ld * D * D * r;
    R tr[D * D], ti[D * D];
#pragma unroll
    for (int xy = 0; xy < D * D; ++xy) {  // xy = x + D y
      const S2 t = th[base + ld * xy];
      tr[xy] = (R)t.x;
      ti[xy] = (R)t.y;
    }
    R lr = (R)1, li = (R)0;
    if (xo) {
      const S2 a = lam[l];
      lr = (R)a.x;
      li = (R)a.y;
    }
#pragma unroll
    for (int pq = 0; pq < D * D; ++pq) {  // pq = p + D q; G[p, q, x, y] at pq + D^2 xy
      R ar = (R)0, ai = (R)0;
#pragma unroll
      for (int xy = 0; xy < D * D; ++xy) {
        const R a = gr[pq + D * D * xy], b = gi[pq + D * D * xy];
        ar = fma(a, tr[xy], fma(-b, ti[xy], ar));
        ai = fma(a, ti[xy], fma(b, tr[xy], ai));
      }
      S2 o;
      o.x = ar;
      o.y = ai;
      to[base + ld * pq] = o;
      if (xo) {
        S2 x;
        x.x = lr * ar - li * ai;
        x.y = lr * ai + li * ar;
        xo[base + ld * pq] = x;
      }
    }
  }
}

template <typename S2, typename R>
static int launch_gate_t(const GateBatch& b, int64_t max_elems, hipStream_t s) {
  const unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>((max_elems + 255) / 256, 4096));
  const dim3 grid(gx, (unsigned)b.count);
  switch (b.d) {
    case 2: hipLaunchKernelGGL((gate_apply_kernel<S2, R, 2>), grid, dim3(256), 0, s, b); break;
    case 3: hipLaunchKernelGGL((gate_apply_kernel<S2, R, 3>), grid, dim3(256), 0, s, b); break;
    case 4: hipLaunchKernelGGL((gate_apply_kernel<S2, R, 4>), grid, dim3(256), 0, s, b); break;
    default: set_error("gate: physical dimension outside 2..4"); return TENET_ERR_UNSUPPORTED;
  }
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

int launch_gate_apply(const GateBatch& b, int64_t max_elems, int dtype, hipStream_t s) {
  if (b.count <= 0 || max_elems <= 0) return TENET_OK;
  return dtype == TENET_C128 ? launch_gate_t<double2, double>(b, max_elems, s) : launch_gate_t<float2, float>(b, max_elems, s);
}

}  // namespace tenet
