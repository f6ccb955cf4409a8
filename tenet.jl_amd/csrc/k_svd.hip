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
#include <hip/hip_ext.h>

#include "../../include/tenet_hip.h"
#include "internal.h"

namespace tenet {

constexpr int SV_THREADS = 256;
// the round kernel: two waves per SIMD, the Gram, the inner solve and the
// rotation split so that each thread carries half the work of a 256-thread
// workgroup's thread
constexpr int SR_THREADS = 512;
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
  union {
    struct {  // row staging (Gram accumulation, rotation of X and V)
      double xr[RS][B2MAX + 1];
      double xi[RS][B2MAX + 1];
    };
    struct {  // second Gram buffer of the inner solve
      double hr[B2MAX][B2MAX + 1];
      double hi[B2MAX][B2MAX + 1];
    };
  };
  double rc[B2MAX / 2], rsr[B2MAX / 2], rsi[B2MAX / 2];
  int rp[B2MAX / 2], rq[B2MAX / 2], ract[B2MAX / 2];
  double rmax[B2MAX / 2];  // per rotation slot: max |g_pq|^2 / (g_pp g_qq) seen
  double rcs[SR_THREADS / 64][B2MAX / 2][4];  // per wave, per slot: c, Re s, Im s, (pad)
  int rpq[SR_THREADS / 64][B2MAX / 2];         // per wave, per slot: rotation active
  int any_outer;
};

__device__ __forceinline__ int pair_col(int l, int I, int J, int b) { return l < b ? I * b + l : J * b + (l - b); }

// Stage rows [r0, r0+rows) of the pair's 2b columns of M (ld) into LDS.
// All of a thread's loads (64 x 32 / 512 = 4) are issued before any is
// written to LDS, so the staging costs one memory latency, not four.
constexpr int STAGE_PER = RS * B2MAX / SR_THREADS;
__device__ __forceinline__ void load_rows(double2 (&v)[STAGE_PER], const double* M, int64_t ld, int r0, int rows,
                                          int I, int J, int b) {
  const int n2 = 2 * b;
  const double2* M2 = reinterpret_cast<const double2*>(M);
#pragma unroll
  for (int u = 0; u < STAGE_PER; ++u) {
    const int e = threadIdx.x + u * SR_THREADS;
    const int r = e % RS, l = e / RS;
    v[u].x = 0.0;
    v[u].y = 0.0;
    if (l < n2 && r < rows) v[u] = M2[(size_t)pair_col(l, I, J, b) * ld + r0 + r];
  }
}

__device__ __forceinline__ void store_rows(SvSmem& sm, const double2 (&v)[STAGE_PER], int n2) {
#pragma unroll
  for (int u = 0; u < STAGE_PER; ++u) {
    const int e = threadIdx.x + u * SR_THREADS;
    const int r = e % RS, l = e / RS;
    if (l < n2) {
      sm.xr[r][l] = v[u].x;
      sm.xi[r][l] = v[u].y;
    }
  }
}

__device__ __forceinline__ void stage_rows(SvSmem& sm, const double* M, int64_t ld, int r0, int rows, int I, int J,
                                           int b) {
  double2 v[STAGE_PER];
  load_rows(v, M, ld, r0, rows, I, J, b);
  store_rows(sm, v, 2 * b);
}

// acc += staged^H staged for this thread's 2x2 block of the Gram matrix:
// slot t' = t % 256 < (n2/2)^2 owns rows {2i, 2i+1} x cols {2j, 2j+1},
// i = t' % (n2/2), j = t' / (n2/2); acc[0..3] = (2i,2j), (2i+1,2j),
// (2i,2j+1), (2i+1,2j+1).  Threads t < 256 take the staged rows [0, 32),
// threads t >= 256 the rows [32, 64); gram_combine adds the two halves.
constexpr int GE_PER_THREAD = 4;
constexpr int GRAM_SLOTS = 256;

__device__ __forceinline__ void gram_accum(SvSmem& sm, int n2, int rows, double* ar, double* ai) {
  const int h = n2 / 2;
  const int t = threadIdx.x % GRAM_SLOTS, half = threadIdx.x / GRAM_SLOTS;
  if (t >= h * h) return;
  const int i0 = 2 * (t % h), j0 = 2 * (t / h);
  const int rlo = half * (RS / 2), rhi = min(rows, rlo + RS / 2);
  double s0r = 0, s0i = 0, s1r = 0, s1i = 0, s2r = 0, s2i = 0, s3r = 0, s3i = 0;
  for (int r = rlo; r < rhi; ++r) {
    const double ar0 = sm.xr[r][i0], ai0 = sm.xi[r][i0];
    const double ar1 = sm.xr[r][i0 + 1], ai1 = sm.xi[r][i0 + 1];
    const double br0 = sm.xr[r][j0], bi0 = sm.xi[r][j0];
    const double br1 = sm.xr[r][j0 + 1], bi1 = sm.xi[r][j0 + 1];
    // conj(a) * b
    s0r += ar0 * br0 + ai0 * bi0;
    s0i += ar0 * bi0 - ai0 * br0;
    s1r += ar1 * br0 + ai1 * bi0;
    s1i += ar1 * bi0 - ai1 * br0;
    s2r += ar0 * br1 + ai0 * bi1;
    s2i += ar0 * bi1 - ai0 * br1;
    s3r += ar1 * br1 + ai1 * bi1;
    s3i += ar1 * bi1 - ai1 * br1;
  }
  ar[0] += s0r;
  ai[0] += s0i;
  ar[1] += s1r;
  ai[1] += s1i;
  ar[2] += s2r;
  ai[2] += s2i;
  ar[3] += s3r;
  ai[3] += s3i;
}

// Threads t < 256 end with the full sums: the upper half's partials pass
// through the (then idle) staging rows.
__device__ __forceinline__ void gram_combine(SvSmem& sm, int n2, double* ar, double* ai) {
  const int h = n2 / 2;
  const int t = threadIdx.x % GRAM_SLOTS, half = threadIdx.x / GRAM_SLOTS;
  double* buf = &sm.xr[0][0];
  __syncthreads();
  if (half == 1 && t < h * h) {
#pragma unroll
    for (int u = 0; u < GE_PER_THREAD; ++u) {
      buf[(2 * u) * GRAM_SLOTS + t] = ar[u];
      buf[(2 * u + 1) * GRAM_SLOTS + t] = ai[u];
    }
  }
  __syncthreads();
  if (half == 0 && t < h * h) {
#pragma unroll
    for (int u = 0; u < GE_PER_THREAD; ++u) {
      ar[u] += buf[(2 * u) * GRAM_SLOTS + t];
      ai[u] += buf[(2 * u + 1) * GRAM_SLOTS + t];
    }
  }
}

// Gram entry index (row, col) of this thread's slot t (0..3)
__device__ __forceinline__ void gram_slot(int n2, int slot, int& row, int& col) {
  const int h = n2 / 2;
  const int t = threadIdx.x;
  row = 2 * (t % h) + (slot & 1);
  col = 2 * (t / h) + (slot >> 1);
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
  gram_combine(sm, n2, ar, ai);
}

// Diagonalise the Hermitian Gram (n2 x n2, n2 <= 32) by one cyclic Jacobi
// sweep in the round-robin (circle) order over all n2 columns: n2 - 1
// steps, step s applies n2/2 disjoint rotations J_k (slot k pairs columns
// p_k < q_k), G <- J^H G J and W <- W J.  (The cross rounds, which pair only
// columns of different blocks, run jacobi_cross32 below.)
//
// One workgroup barrier per step, with the rotations formed one step ahead.
// G is double-buffered (the second buffer lies over the row-staging area,
// idle during the solve) and only its upper triangle is kept current
// (entries below it read as conjugated mirrors).  In the circle order the
// columns of slot k at step s + 1 come from two different slots k1, k2 of
// step s, so the entries step s + 1's rotation k needs are
//   G'[p'][p'], G'[q'][q'] = new diagonals of rotations k1, k2 of step s,
//   G'[p'][q']             = one entry of J_k1^H B J_k2, B the 2x2 block
//                            {p_k1, q_k1} x {p_k2, q_k2} of G before step s,
// all known from G_s and step s's rotations.  Wave 7 (one lane per slot)
// writes step s's diagonal blocks (closed form) and forms step s + 1's
// rotations while waves 0-1 update the off-diagonal 2 x 2 blocks (one thread
// per block: both output rows) and waves 2-6 rotate W with step s's
// rotations; the rotations (c, s and the new diagonal) pass through LDS
// slots double-buffered by step parity.
// columns of slot k at step s of the circle order
__device__ __forceinline__ void circle_pair(int k, int step, int n2, int& p, int& q) {
  // circle_pos without the modulo: i - 1 + r < 2 (n2 - 1)
  const int m = n2 - 1;
  const int i0 = k, i1 = n2 - 1 - k;
  int a0 = i0 - 1 + step, a1 = i1 - 1 + step;
  a0 = a0 >= m ? a0 - m : a0;
  a1 = a1 >= m ? a1 - m : a1;
  p = i0 == 0 ? 0 : 1 + a0;
  q = i1 == 0 ? 0 : 1 + a1;
  if (p > q) {
    const int t = p;
    p = q;
    q = t;
  }
}

typedef double GramRow[B2MAX + 1];

// 1/sqrt(x) for x > 0: hardware estimate and two Newton steps in correction
// form y <- y + y (1 - x y^2) / 2 (each squares the relative error; the
// estimate is good to ~2^-23, so two steps reach double precision).
__device__ __forceinline__ double rsqrt_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  y = fma(0.5 * y, fma(-x * y, y, 1.0), y);
  y = fma(0.5 * y, fma(-x * y, y, 1.0), y);
  return y;
}

// 1/x for x > 0: hardware estimate and two Newton steps y <- y + y (1 - x y)
__device__ __forceinline__ double rcp_nr(double x) {
  double y = __builtin_amdgcn_rcp(x);
  y = fma(y, fma(-x, y, 1.0), y);
  y = fma(y, fma(-x, y, 1.0), y);
  return y;
}

// Rotation zeroing G[p][q] of [[a, g], [conj(g), d]].  With
// delta = (d - a) / 2 and r = sqrt(delta^2 + |g|^2) the angle has
// cos 2theta = |delta| / r and sin 2theta = |g| / r, so c = cos theta =
// sqrt((1 + |delta| / r) / 2) and s = sign(delta) g / (2 r c) (the
// smaller rotation of Rutishauser's formulas, tan theta = |g| / (|delta| + r)).
// The new diagonal is a - sign(delta) t |g|, d + sign(delta) t |g| with
// t |g| = |g|^2 / (|delta| + r) = |g|^2 / (2 r c^2).
// Identity when |g| <= tol sqrt(a d) or |g| <= abs_tol (compared squared).
// x of lane (l + 1) mod 16 within each 16-lane row (DPP row_ror:15, on the
// VALU; lanes 0-15 hold the slots of the cross-block inner solve)
__device__ __forceinline__ double from_next_lane(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), 0x12F, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x12F, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

struct Rot {
  double c, sr, si;  // J = [[c, s], [-conj(s), c]]
  double an, dn;     // new diagonal
  double rel2;       // |g|^2 / (a d), for the convergence test
  int act;
};

__device__ __forceinline__ Rot form_rotation(double ra, double rd, double rxr, double rxi, double tol2, double abs2) {
  Rot o;
  const double g2 = rxr * rxr + rxi * rxi;
  const double ad = fmax(ra, 0.0) * fmax(rd, 0.0);
  o.act = (g2 > tol2 * ad && g2 > abs2) ? 1 : 0;
  o.c = 1.0;
  o.sr = 0.0;
  o.si = 0.0;
  o.an = ra;
  o.dn = rd;
  if (o.act) {
    const double dl = 0.5 * (rd - ra);
    const double ir = rsqrt_nr(fma(dl, dl, g2));    // 1 / r
    const double v = fma(0.5 * fabs(dl), ir, 0.5);  // c^2
    const double y = rsqrt_nr(v);                   // 1 / c
    o.c = v * y;
    const double sn = 0.5 * ir * y;
    const double tg = (0.5 * g2) * (ir * (y * y));
    const bool pos = dl >= 0.0;
    o.sr = pos ? sn * rxr : -(sn * rxr);
    o.si = pos ? sn * rxi : -(sn * rxi);
    o.an = pos ? ra - tg : ra + tg;
    o.dn = pos ? rd + tg : rd - tg;
  }
  o.rel2 = ad > 0.0 ? g2 * rcp_nr(ad) : 0.0;
  return o;
}

// row h of B' = J_k^H B J_l for the 2x2 block B = [B0; B1] and the rotations
// J = [[c, e], [-conj(e), c]]: row h of T = J_k^H B is a B0 + b B1 with
// (a, b) = (c_k, -e_k) for h = 0 and (conj(e_k), c_k) for h = 1, and row h of
// B' is (c_l T0 - conj(e_l) T1, e_l T0 + c_l T1).
__device__ __forceinline__ void block_row(bool h1, double ck, double ekr, double eki, double cl, double elr,
                                          double eli, const double (&b)[8], double& n0r, double& n0i, double& n1r,
                                          double& n1i) {
  const double a_r = h1 ? ekr : ck, a_i = h1 ? -eki : 0.0;
  const double b_r = h1 ? ck : -ekr, b_i = h1 ? 0.0 : -eki;
  // b = {B00 re, im, B01 re, im, B10 re, im, B11 re, im}
  const double t0r = (a_r * b[0] - a_i * b[1]) + (b_r * b[4] - b_i * b[5]);
  const double t0i = (a_r * b[1] + a_i * b[0]) + (b_r * b[5] + b_i * b[4]);
  const double t1r = (a_r * b[2] - a_i * b[3]) + (b_r * b[6] - b_i * b[7]);
  const double t1i = (a_r * b[3] + a_i * b[2]) + (b_r * b[7] + b_i * b[6]);
  n0r = cl * t0r - (elr * t1r + eli * t1i);
  n0i = cl * t0i - (elr * t1i - eli * t1r);
  n1r = (elr * t0r - eli * t0i) + cl * t1r;
  n1i = (elr * t0i + eli * t0r) + cl * t1i;
}

// G entry (r, c), r != c, from the upper triangle (an entry below it is the
// conjugate of its mirror), and the matching store
__device__ __forceinline__ void g_get(const GramRow* gr, const GramRow* gi, int r, int c, double& re, double& im) {
  const int lo = min(r, c), hi = max(r, c);
  re = gr[lo][hi];
  im = r < c ? gi[lo][hi] : -gi[lo][hi];
}
__device__ __forceinline__ void g_put(GramRow* gr, GramRow* gi, int r, int c, double re, double im) {
  const int lo = min(r, c), hi = max(r, c);
  gr[lo][hi] = re;
  gi[lo][hi] = r < c ? im : -im;
}

// Slot and side (0: p, 1: q) of column c at step s of the circle order over
// n2 columns (the inverse of circle_pair: the column at position i >= 1 is
// 1 + (i - 1 + s) mod (n2 - 1), position i belongs to slot min(i, n2 - 1 - i))
__device__ __forceinline__ void circle_slot(int c, int step, int n2, int& slot, int& side) {
  const int m = n2 - 1;
  int i = 0;
  if (c != 0) {
    int t = c - 1 - step;
    t = t < 0 ? t + m : t;
    i = 1 + t;
  }
  slot = min(i, n2 - 1 - i);
  int p, q;
  circle_pair(slot, step, n2, p, q);
  side = c == q ? 1 : 0;
}

constexpr int IROT_WAVE = 7;      // rotation wave of the full inner sweep
constexpr int IW_FIRST = 2 * 64;  // first W thread (waves 2-6)
constexpr int IW_THREADS = 5 * 64;

__device__ __forceinline__ void jacobi_inner(SvSmem& sm, int n2, double tol, double abs_tol) {
  const int tid = threadIdx.x, wave = tid >> 6, ln = tid & 63;
  const int npair = n2 / 2, nsteps = n2 - 1;
  // W = I on the pair's n2 columns and 0 elsewhere in the 32 x 32 block (the
  // MFMA rotation runs over all 32 columns)
  for (int e = tid; e < B2MAX * B2MAX; e += SR_THREADS) {
    const int i = e % B2MAX, c = e / B2MAX;
    sm.wr[i][c] = (i == c && i < n2) ? 1.0 : 0.0;
    sm.wi[i][c] = 0.0;
  }
  if (tid == 0) sm.any_outer = 0;
  const double tol2 = tol * tol, abs2 = abs_tol * abs_tol;
  const bool gwave = wave < 2, wwave = tid >= IW_FIRST && tid < IW_FIRST + IW_THREADS, rwave = wave == IROT_WAVE;
  // G: the npair (npair - 1) / 2 blocks bk < bl in row-major order, one thread each
  int bk = 0, bl = 1;
  bool gown = false;
  if (gwave) {
    gown = tid < npair * (npair - 1) / 2;
    int rem = gown ? tid : 0, k = 0;
    while (k < npair - 1 && rem >= npair - 1 - k) {
      rem -= npair - 1 - k;
      ++k;
    }
    bk = k;
    bl = k + 1 + rem;
  }
  // W: element e = (tid - IW_FIRST) + IW_THREADS u is (row e / npair, slot e % npair); npair n2 <= 512
  int wrow[2], wslot[2];
  bool wown[2];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int e = (tid - IW_FIRST) + IW_THREADS * u;
    wown[u] = wwave && e < npair * n2;
    const int ec = wown[u] ? e : 0;
    wrow[u] = ec / npair;
    wslot[u] = ec % npair;
  }
  // rotation wave: lane k < npair owns slot k
  const bool rown = rwave && ln < npair;
  const int k = rwave ? min(ln, npair - 1) : 0;
  Rot rt;
  double rgr = 0.0, rgi = 0.0;  // G[p_k][q_k] the current rotation was formed from
  double rmax2 = 0.0;
  int any = 0;
  __syncthreads();
  if (rwave) {
    int p, q;
    circle_pair(k, 0, n2, p, q);
    rgr = sm.gr[p][q];
    rgi = sm.gi[p][q];
    rt = form_rotation(sm.gr[p][p], sm.gr[q][q], rgr, rgi, tol2, abs2);
    if (rown) {
      sm.rcs[0][k][0] = rt.c;
      sm.rcs[0][k][1] = rt.sr;
      sm.rcs[0][k][2] = rt.si;
      sm.rcs[2][k][0] = rt.an;
      sm.rcs[2][k][1] = rt.dn;
      sm.rpq[0][k] = rt.act;
      rmax2 = fmax(rmax2, rt.rel2);
      any |= rt.act;
    }
  }
  __syncthreads();
  for (int step = 0; step < nsteps; ++step) {
    const int par = step & 1;
    const GramRow* cr = par ? sm.hr : sm.gr;
    const GramRow* ci = par ? sm.hi : sm.gi;
    GramRow* nr = par ? sm.gr : sm.hr;
    GramRow* ni = par ? sm.gi : sm.hi;
    if (gwave) {
      if (gown) {
        int p0, q0, p1, q1;
        circle_pair(bk, step, n2, p0, q0);
        circle_pair(bl, step, n2, p1, q1);
        double b[8];
        g_get(cr, ci, p0, p1, b[0], b[1]);
        g_get(cr, ci, p0, q1, b[2], b[3]);
        g_get(cr, ci, q0, p1, b[4], b[5]);
        g_get(cr, ci, q0, q1, b[6], b[7]);
        const double ck = sm.rcs[par][bk][0], ekr = sm.rcs[par][bk][1], eki = sm.rcs[par][bk][2];
        const double cl = sm.rcs[par][bl][0], elr = sm.rcs[par][bl][1], eli = sm.rcs[par][bl][2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          double n0r, n0i, n1r, n1i;
          block_row(h != 0, ck, ekr, eki, cl, elr, eli, b, n0r, n0i, n1r, n1i);
          const int ph = h ? q0 : p0;
          g_put(nr, ni, ph, p1, n0r, n0i);
          g_put(nr, ni, ph, q1, n1r, n1i);
        }
      }
    } else if (wwave) {
      // W <- W J on this thread's elements
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int ks = wslot[u];
        if (!wown[u] || !sm.rpq[par][ks]) continue;
        int wp, wq;
        circle_pair(ks, step, n2, wp, wq);
        const int row = wrow[u];
        const double pr = sm.wr[row][wp], pi = sm.wi[row][wp], qr = sm.wr[row][wq], qi = sm.wi[row][wq];
        const double cw = sm.rcs[par][ks][0], er = sm.rcs[par][ks][1], ei = sm.rcs[par][ks][2];
        sm.wr[row][wp] = cw * pr - (er * qr + ei * qi);
        sm.wi[row][wp] = cw * pi - (er * qi - ei * qr);
        sm.wr[row][wq] = (er * pr - ei * pi) + cw * qr;
        sm.wi[row][wq] = (er * pi + ei * pr) + cw * qi;
      }
    } else if (rwave) {
      // this step's diagonal blocks, then the next step's rotations
      if (rown) {
        int p, q;
        circle_pair(k, step, n2, p, q);
        nr[p][p] = rt.an;
        ni[p][p] = 0.0;
        nr[q][q] = rt.dn;
        ni[q][q] = 0.0;
        nr[p][q] = rt.act ? 0.0 : rgr;
        ni[p][q] = rt.act ? 0.0 : rgi;
      }
      if (step + 1 < nsteps) {
        int pn, qn, k1, h1, k2, h2, pk1, qk1, pk2, qk2;
        circle_pair(k, step + 1, n2, pn, qn);
        circle_slot(pn, step, n2, k1, h1);
        circle_slot(qn, step, n2, k2, h2);
        circle_pair(k1, step, n2, pk1, qk1);
        circle_pair(k2, step, n2, pk2, qk2);
        double b[8];
        g_get(cr, ci, pk1, pk2, b[0], b[1]);
        g_get(cr, ci, pk1, qk2, b[2], b[3]);
        g_get(cr, ci, qk1, pk2, b[4], b[5]);
        g_get(cr, ci, qk1, qk2, b[6], b[7]);
        const double c1 = sm.rcs[par][k1][0], s1r = sm.rcs[par][k1][1], s1i = sm.rcs[par][k1][2];
        const double c2 = sm.rcs[par][k2][0], s2r = sm.rcs[par][k2][1], s2i = sm.rcs[par][k2][2];
        const double an = sm.rcs[2 + par][k1][h1], dn = sm.rcs[2 + par][k2][h2];
        double n0r, n0i, n1r, n1i;
        block_row(h1 != 0, c1, s1r, s1i, c2, s2r, s2i, b, n0r, n0i, n1r, n1i);
        rgr = h2 ? n1r : n0r;
        rgi = h2 ? n1i : n0i;
        rt = form_rotation(an, dn, rgr, rgi, tol2, abs2);
        if (rown) {
          sm.rcs[par ^ 1][k][0] = rt.c;
          sm.rcs[par ^ 1][k][1] = rt.sr;
          sm.rcs[par ^ 1][k][2] = rt.si;
          sm.rcs[2 + (par ^ 1)][k][0] = rt.an;
          sm.rcs[2 + (par ^ 1)][k][1] = rt.dn;
          sm.rpq[par ^ 1][k] = rt.act;
          rmax2 = fmax(rmax2, rt.rel2);
          any |= rt.act;
        }
      }
    }
    __syncthreads();
  }
  if (rown) sm.rmax[k] = rmax2;
  if (rown && any) sm.any_outer = 1;
  __syncthreads();
}

// Cross-block inner solve for 2b = 32 (every cross round: a matrix with more
// than one block pair has b = 16) with the rotations formed one step ahead.
// In the cross order p_k = k at every step and q_k(s + 1) = q_{k+1}(s)
// (slots mod 16), so the Gram entries rotation k of step s + 1 needs are
//   G'[k][k]       = the new diagonal a'_k of rotation k of step s,
//   G'[q][q]       = the new diagonal d'_{k+1} of rotation k + 1 of step s,
//   G'[k][q_{k+1}] = row p_k, column q_{k+1} of J_k^H B J_{k+1}, B the 2x2
//                    block {k, q_k} x {k+1, q_{k+1}} of G before step s,
// all known from G_s and step s's rotations.  Wave 7 (lanes 0-15: one slot
// each) forms them and step s + 1's rotations while 120 threads of waves
// 0-1 update the G blocks above the diagonal (one 2 x 2 block each, read
// once; only G's upper triangle is kept, entries below it read as
// conjugated mirrors) and waves 4-6 rotate W, which
// they keep in registers, with step s's rotations; the rotations pass
// through LDS slots double-buffered by step parity, so a step still takes
// one workgroup barrier.  Wave 7 also writes the diagonal blocks (closed
// form).
constexpr int XB = 16;         // b
constexpr int XROT_WAVE = 7;   // rotation wave

__device__ __forceinline__ void jacobi_cross32(SvSmem& sm, double tol, double abs_tol) {
  const int tid = threadIdx.x, wave = tid >> 6, ln = tid & 63;
  if (tid == 0) sm.any_outer = 0;
  const double tol2 = tol * tol, abs2 = abs_tol * abs_tol;
  // roles (wave-uniform)
  const bool gwave = wave < 4, wwave = wave >= 4 && wave < XROT_WAVE, rwave = wave == XROT_WAVE;
  // G: blocks bk < bl in row-major order, one thread each (both output rows:
  // the 2 x 2 block is read once)
  int bk = 0, bl = 1;
  bool gown = false;
  if (gwave) {
    const int u = tid;
    gown = u < XB * (XB - 1) / 2;
    int rem = gown ? u : 0, k = 0;
    while (k < XB - 1 && rem >= XB - 1 - k) {
      rem -= XB - 1 - k;
      ++k;
    }
    bk = k;
    bl = k + 1 + rem;
  }
  // rotation wave: lane k < 16 owns slot k
  const int k = ln & (XB - 1), kn = (k + 1) & (XB - 1);
  // W lives in the registers of waves 4-6: lane k of 16-lane row group wg
  // (0..11) holds, for rows wg, wg + 12 and wg + 24 (< 32), P = W[row][k]
  // (slot k's p column is k at every step) and Q = W[row][q_k(step)] with
  // q_k(s) = 16 + (k + s) mod 16; since q_k(s + 1) = q_{k+1}(s), Q moves
  // one lane down between steps (a DPP row rotation), and W goes to LDS once
  // at the end.
  const int wg = wwave ? ((tid - 4 * 64) >> 4) : 0;
  double Pr[3], Pi[3], Qr[3], Qi[3];
#pragma unroll
  for (int u = 0; u < 3; ++u) {
    const int row = wg + 12 * u;
    Pr[u] = row == k ? 1.0 : 0.0;
    Pi[u] = 0.0;
    Qr[u] = row == XB + k ? 1.0 : 0.0;
    Qi[u] = 0.0;
  }
  const bool rown = rwave && ln < XB;
  Rot rt;
  double gr = 0.0, gi = 0.0;  // G[k][q_k] the current rotation was formed from
  double rmax2 = 0.0;
  int any = 0;
  if (rwave) {
    const int q = XB + k;
    gr = sm.gr[k][q];
    gi = sm.gi[k][q];
    rt = form_rotation(sm.gr[k][k], sm.gr[q][q], gr, gi, tol2, abs2);
    if (rown) {
      sm.rcs[0][k][0] = rt.c;
      sm.rcs[0][k][1] = rt.sr;
      sm.rcs[0][k][2] = rt.si;
      sm.rpq[0][k] = rt.act;
      rmax2 = fmax(rmax2, rt.rel2);
      any |= rt.act;
    }
  }
  __syncthreads();
#pragma unroll 2
  for (int step = 0; step < XB; ++step) {
    const int par = step & 1;
    const GramRow* cr = par ? sm.hr : sm.gr;
    const GramRow* ci = par ? sm.hi : sm.gi;
    GramRow* nr = par ? sm.gr : sm.hr;
    GramRow* ni = par ? sm.gi : sm.hi;
    if (gwave) {
      // only the upper triangle of G is kept up to date: an entry below the
      // diagonal is read as the conjugate of its mirror
      const int p0 = bk, q0 = XB + ((bk + step) & (XB - 1));
      const int p1 = bl, q1 = XB + ((bl + step) & (XB - 1));
      const int qa = min(q0, q1), qb = max(q0, q1);
      const double qs = q0 <= q1 ? 1.0 : -1.0;
      double b[8];
      b[0] = cr[p0][p1];
      b[1] = ci[p0][p1];
      b[2] = cr[p0][q1];
      b[3] = ci[p0][q1];
      b[4] = cr[p1][q0];
      b[5] = -ci[p1][q0];
      b[6] = cr[qa][qb];
      b[7] = qs * ci[qa][qb];
      const double ck = sm.rcs[par][bk][0], ekr = sm.rcs[par][bk][1], eki = sm.rcs[par][bk][2];
      const double cl = sm.rcs[par][bl][0], elr = sm.rcs[par][bl][1], eli = sm.rcs[par][bl][2];
      if (gown) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          double n0r, n0i, n1r, n1i;
          block_row(h != 0, ck, ekr, eki, cl, elr, eli, b, n0r, n0i, n1r, n1i);
          if (h == 0) {
            nr[p0][p1] = n0r;
            ni[p0][p1] = n0i;
            nr[p0][q1] = n1r;
            ni[p0][q1] = n1i;
          } else {
            nr[p1][q0] = n0r;
            ni[p1][q0] = -n0i;
            nr[qa][qb] = n1r;
            ni[qa][qb] = qs * n1i;
          }
        }
      }
    } else if (wwave) {
      if (step > 0) {
#pragma unroll
        for (int u = 0; u < 3; ++u) {
          Qr[u] = from_next_lane(Qr[u]);
          Qi[u] = from_next_lane(Qi[u]);
        }
      }
      if (sm.rpq[par][k]) {
        const double c = sm.rcs[par][k][0], er = sm.rcs[par][k][1], ei = sm.rcs[par][k][2];
#pragma unroll
        for (int u = 0; u < 3; ++u) {
          const double pr = Pr[u], pi = Pi[u], qr = Qr[u], qi = Qi[u];
          Pr[u] = c * pr - (er * qr + ei * qi);
          Pi[u] = c * pi - (er * qi - ei * qr);
          Qr[u] = (er * pr - ei * pi) + c * qr;
          Qi[u] = (er * pi + ei * pr) + c * qi;
        }
      }
    } else {
      // rotation wave: this step's diagonal blocks, then the next step's rotations
      const int q = XB + ((k + step) & (XB - 1));
      const int qn = XB + ((kn + step) & (XB - 1));
      const int ka = min(k, kn), kb2 = max(k, kn), qa = min(q, qn), qb = max(q, qn);
      const double ks = k <= kn ? 1.0 : -1.0, qs = q <= qn ? 1.0 : -1.0;
      double b[8];
      b[0] = cr[ka][kb2];
      b[1] = ks * ci[ka][kb2];
      b[2] = cr[k][qn];
      b[3] = ci[k][qn];
      b[4] = cr[kn][q];
      b[5] = -ci[kn][q];
      b[6] = cr[qa][qb];
      b[7] = qs * ci[qa][qb];
      if (rown) {
        const double ogr = rt.act ? 0.0 : gr, ogi = rt.act ? 0.0 : gi;
        nr[k][k] = rt.an;
        ni[k][k] = 0.0;
        nr[q][q] = rt.dn;
        ni[q][q] = 0.0;
        nr[k][q] = ogr;
        ni[k][q] = ogi;
      }
      if (step + 1 < XB) {
        // slot k + 1's rotation and new diagonal d' (lanes 0-15 hold slots 0-15)
        const double cn = from_next_lane(rt.c), srn = from_next_lane(rt.sr), sin_ = from_next_lane(rt.si);
        const double dnn = from_next_lane(rt.dn);
        double n0r, n0i, n1r, n1i;
        block_row(false, rt.c, rt.sr, rt.si, cn, srn, sin_, b, n0r, n0i, n1r, n1i);
        gr = n1r;
        gi = n1i;
        rt = form_rotation(rt.an, dnn, gr, gi, tol2, abs2);
        if (rown) {
          sm.rcs[par ^ 1][k][0] = rt.c;
          sm.rcs[par ^ 1][k][1] = rt.sr;
          sm.rcs[par ^ 1][k][2] = rt.si;
          sm.rpq[par ^ 1][k] = rt.act;
          rmax2 = fmax(rmax2, rt.rel2);
          any |= rt.act;
        }
      }
    }
    __syncthreads();
  }
  if (wwave) {
    const int q = XB + ((k + XB - 1) & (XB - 1));  // q_k of the last step
#pragma unroll
    for (int u = 0; u < 3; ++u) {
      const int row = wg + 12 * u;
      if (row < 2 * XB) {
        sm.wr[row][k] = Pr[u];
        sm.wi[row][k] = Pi[u];
        sm.wr[row][q] = Qr[u];
        sm.wi[row][q] = Qi[u];
      }
    }
  }
  if (rown) sm.rmax[k] = rmax2;
  if (rown && any) sm.any_outer = 1;
  __syncthreads();
}

// The staged rows [rb, rb + rows) of the pair's columns of M <- M W.  Thread
// t: row t % 64 of the staged chunk and 4 consecutive output columns
// (group t / 64).
constexpr int APPLY_COLS = SR_THREADS / RS > 0 ? B2MAX / (SR_THREADS / RS) : B2MAX;
__device__ __forceinline__ void apply_staged(SvSmem& sm, double* M, int64_t ld, int rb, int rows, int I, int J,
                                             int b) {
  const int n2 = 2 * b;
  const int r = threadIdx.x % RS, cg = APPLY_COLS * (threadIdx.x / RS);
  if (r < rows && cg < n2) {
    double sr[APPLY_COLS], si[APPLY_COLS];
#pragma unroll
    for (int j = 0; j < APPLY_COLS; ++j) sr[j] = si[j] = 0.0;
    for (int t = 0; t < n2; ++t) {
      const double xr = sm.xr[r][t], xi = sm.xi[r][t];
#pragma unroll
      for (int j = 0; j < APPLY_COLS; ++j) {
        const double wr = sm.wr[t][cg + j], wi = sm.wi[t][cg + j];
        sr[j] += xr * wr - xi * wi;
        si[j] += xr * wi + xi * wr;
      }
    }
#pragma unroll
    for (int j = 0; j < APPLY_COLS; ++j) {
      if (cg + j < n2) {
        double2 v;
        v.x = sr[j];
        v.y = si[j];
        reinterpret_cast<double2*>(M)[(size_t)pair_col(cg + j, I, J, b) * ld + rb + r] = v;
      }
    }
  }
}

// rows [r0, r1) of the pair's columns of M <- M W, one staged chunk at a time
__device__ void apply_rows(SvSmem& sm, double* M, int64_t ld, int r0, int r1, int I, int J, int b) {
  for (int rb = r0; rb < r1; rb += RS) {
    const int rows = min(RS, r1 - rb);
    __syncthreads();
    stage_rows(sm, M, ld, rb, rows, I, J, b);
    __syncthreads();
    apply_staged(sm, M, ld, rb, rows, I, J, b);
  }
  __syncthreads();
}

// ---- FP64 MFMA forms of the Gram and of the rotation (v_mfma_f64_16x16x4_f64)
//
// Operand layout of the 16x16x4 FP64 product (lane l): A[i = l % 16][k = l / 16],
// B[k = l / 16][j = l % 16]; accumulator register r of lane l holds
// D[i = 4 r + l / 16][j = l % 16].  The pair's 2b <= 32 columns are two
// 16-column blocks (columns >= 2b staged as zeros).
typedef double svd_d4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ svd_d4v mfma_f64(double a, double b, svd_d4v c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// All 32 staged columns (zeros beyond 2b and below `rows`), so that the MFMA
// operands never read stale LDS.
__device__ __forceinline__ void store_rows_all(SvSmem& sm, const double2 (&v)[STAGE_PER]) {
#pragma unroll
  for (int u = 0; u < STAGE_PER; ++u) {
    const int e = threadIdx.x + u * SR_THREADS;
    const int r = e % RS, l = e / RS;
    sm.xr[r][l] = v[u].x;
    sm.xi[r][l] = v[u].y;
  }
}

// Gram G = X^H X of the pair's columns over rows [r0, r1) into sm.gr / sm.gi.
// With P = Xr^T Xi:  Re G = Xr^T Xr + Xi^T Xi (symmetric), Im G = P - P^T.
// For a 4-row K step the A operand of X^T and the B operand of X are the same
// register (X[k = l / 16][c = l % 16]).  Every block takes fewer real
// products than the four of a complex one (errors O(eps |x_p| |x_q|), the
// order of a plain dot product's): a diagonal block two, S^T S with
// S = Xr + Xi and P, since Re G = S^T S - P - P^T; the off-diagonal block
// three (Gauss): T1 = Xr0^T Xr1, T2 = Xi0^T Xi1 and
// T3 = (Xr0 + Xi0)^T (Xr1 - Xi1) give Re G(0,1) = T1 + T2 and
// Im G(0,1) = T1 - T2 - T3.  Each 64-row chunk is cut into four 16-row
// groups; group g belongs to waves g and g + 4 (the two waves of one SIMD):
// wave g sums S0^T S0, S1^T S1 and P(0,0), wave g + 4 P(1,1), T1, T2 and T3
// (seven products per K step on the SIMD).  The groups' partials are added
// in LDS in a fixed order (1 into 0 and 3 into 2, then 2 into 0:
// deterministic).
__device__ void gram_mfma(SvSmem& sm, const double* X, int64_t ld, int r0, int r1, int I, int J, int b) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int grp = wave & 3;
  const bool hiw = wave >= 4;
  const int kr = lane >> 4, kc = lane & 15;
  svd_d4v a0 = {0, 0, 0, 0}, a1 = {0, 0, 0, 0}, a2 = {0, 0, 0, 0}, a3 = {0, 0, 0, 0};
  double2 v[STAGE_PER];
  load_rows(v, X, ld, r0, min(RS, r1 - r0), I, J, b);
  for (int rb = r0; rb < r1; rb += RS) {
    __syncthreads();
    store_rows_all(sm, v);
    __syncthreads();
    if (rb + RS < r1) load_rows(v, X, ld, rb + RS, min(RS, r1 - rb - RS), I, J, b);
#pragma unroll
    for (int s = 0; s < RS / 16; ++s) {
      const int row = grp * (RS / 4) + 4 * s + kr;
      const double xr0 = sm.xr[row][kc], xi0 = sm.xi[row][kc];
      const double xr1 = sm.xr[row][16 + kc], xi1 = sm.xi[row][16 + kc];
      if (!hiw) {
        const double s0 = xr0 + xi0, s1 = xr1 + xi1;
        a0 = mfma_f64(s0, s0, a0);
        a1 = mfma_f64(s1, s1, a1);
        a2 = mfma_f64(xr0, xi0, a2);
      } else {
        a0 = mfma_f64(xr1, xi1, a0);
        a1 = mfma_f64(xr0, xr1, a1);
        a2 = mfma_f64(xi0, xi1, a2);
        a3 = mfma_f64(xr0 + xi0, xr1 - xi1, a3);
      }
    }
  }
  // partials through the staging area: slot (pair, wave half) of 4 x 4 x 64
  double* red = &sm.xr[0][0];
  const int slot = (grp >> 1) * 2 + (hiw ? 1 : 0);
  __syncthreads();
  if (grp & 1) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      red[((slot * 4 + 0) * 4 + r) * 64 + lane] = a0[r];
      red[((slot * 4 + 1) * 4 + r) * 64 + lane] = a1[r];
      red[((slot * 4 + 2) * 4 + r) * 64 + lane] = a2[r];
      red[((slot * 4 + 3) * 4 + r) * 64 + lane] = a3[r];
    }
  }
  __syncthreads();
  if (!(grp & 1)) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      a0[r] += red[((slot * 4 + 0) * 4 + r) * 64 + lane];
      a1[r] += red[((slot * 4 + 1) * 4 + r) * 64 + lane];
      a2[r] += red[((slot * 4 + 2) * 4 + r) * 64 + lane];
      a3[r] += red[((slot * 4 + 3) * 4 + r) * 64 + lane];
    }
  }
  __syncthreads();
  if (grp == 2) {
    const int s2 = hiw ? 1 : 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      red[((s2 * 4 + 0) * 4 + r) * 64 + lane] = a0[r];
      red[((s2 * 4 + 1) * 4 + r) * 64 + lane] = a1[r];
      red[((s2 * 4 + 2) * 4 + r) * 64 + lane] = a2[r];
      red[((s2 * 4 + 3) * 4 + r) * 64 + lane] = a3[r];
    }
  }
  __syncthreads();
  if (grp == 0) {
    const int s2 = hiw ? 1 : 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      a0[r] += red[((s2 * 4 + 0) * 4 + r) * 64 + lane];
      a1[r] += red[((s2 * 4 + 1) * 4 + r) * 64 + lane];
      a2[r] += red[((s2 * 4 + 2) * 4 + r) * 64 + lane];
      a3[r] += red[((s2 * 4 + 3) * 4 + r) * 64 + lane];
    }
    // S^T S of the diagonal blocks and Re G(0,1) (mirrored) into gr, P into
    // wr (scratch until the inner solve sets W)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = 4 * r + kr, j = kc;
      if (!hiw) {
        sm.gr[i][j] = a0[r];
        sm.gr[16 + i][16 + j] = a1[r];
        sm.wr[i][j] = a2[r];
      } else {
        // the off-diagonal block of Im G goes into wr above the diagonal
        // (zeros below), so that gi = wr - wr^T holds it and its mirror
        sm.wr[16 + i][16 + j] = a0[r];
        sm.gr[i][16 + j] = a1[r] + a2[r];
        sm.gr[16 + j][i] = a1[r] + a2[r];
        sm.wr[i][16 + j] = a1[r] - a2[r] - a3[r];
        sm.wr[16 + j][i] = 0.0;
      }
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < B2MAX * B2MAX; e += SR_THREADS) {
    const int i = e & (B2MAX - 1), j = e / B2MAX;
    const double pij = sm.wr[i][j], pji = sm.wr[j][i];
    sm.gi[i][j] = pij - pji;
    if ((i < 16) == (j < 16)) sm.gr[i][j] -= pij + pji;  // Re G = S^T S - P - P^T
  }
  __syncthreads();
}

// Rows [r0, r1) of the pair's columns of M <- M W by MFMA, with no LDS
// staging: wave w takes the 16-row strips w, w + 8, ... and computes all 32
// output columns of its strip (two 16-column blocks over K = 32, three real
// products each, see below).  Its A operands come straight from memory (lane
// l: row l % 16 of the strip, columns 4s + l / 16: 256 contiguous bytes per
// column and load), W's from LDS, and the outputs go out from the
// accumulators (lanes l, l + 16, l + 32, l + 48: four consecutive rows of one
// column).  A wave reads and writes only its own rows, so the waves run
// without a barrier.
__device__ void apply_mfma(SvSmem& sm, double* M, int64_t ld, int r0, int r1, int I, int J, int b) {
  const int n2 = 2 * b;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int kr = lane >> 4, kc = lane & 15;
  double2* M2 = reinterpret_cast<double2*>(M);
  const size_t safe = (size_t)pair_col(0, I, J, b) * ld + r0;  // a valid element (masked loads read it)
  for (int rs = r0 + 16 * wave; rs < r1; rs += 16 * (SR_THREADS / 64)) {
    const int row = rs + kc;
    double2 x[B2MAX / 4];
#pragma unroll
    for (int s = 0; s < B2MAX / 4; ++s) {
      const int l = 4 * s + kr;
      const bool ok = row < r1 && l < n2;
      const double2 v = M2[ok ? (size_t)pair_col(l, I, J, b) * ld + row : safe];
      x[s].x = ok ? v.x : 0.0;
      x[s].y = ok ? v.y : 0.0;
    }
    // three real products per complex one (Gauss):
    //   T1 = Xr Wr, T2 = Xi Wi, T3 = (Xr + Xi)(Wr + Wi);
    //   Re Y = T1 - T2, Im Y = T3 - T1 - T2
    // (W is unitary, so the cancellation in Im Y costs O(eps |X|))
    svd_d4v t[2][3];
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
#pragma unroll
      for (int u = 0; u < 3; ++u) t[cb][u] = svd_d4v{0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < B2MAX / 4; ++s) {
      const double xr = x[s].x, xi = x[s].y, xs = xr + xi;
#pragma unroll
      for (int cb = 0; cb < 2; ++cb) {
        const double wre = sm.wr[4 * s + kr][16 * cb + kc], wim = sm.wi[4 * s + kr][16 * cb + kc];
        t[cb][0] = mfma_f64(xr, wre, t[cb][0]);
        t[cb][1] = mfma_f64(xi, wim, t[cb][1]);
        t[cb][2] = mfma_f64(xs, wre + wim, t[cb][2]);
      }
    }
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
      const int l = 16 * cb + kc;
      if (l < n2) {
        double2* col = M2 + (size_t)pair_col(l, I, J, b) * ld + rs;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int rr = 4 * r + kr;
          if (rs + rr < r1) {
            double2 o;
            o.x = t[cb][0][r] - t[cb][1][r];
            o.y = t[cb][2][r] - t[cb][0][r] - t[cb][1][r];
            col[rr] = o;
          }
        }
      }
    }
  }
  __syncthreads();
}

// Pair (I, J) of item p in round `round` (mode 0: the fixed pairs (2p, 2p+1);
// mode 1: the circle pairs).
__device__ __forceinline__ void round_pair(int p, int round, int mode, int nb, int& I, int& J) {
  I = 2 * p;
  J = 2 * p + 1;
  if (mode == 1) {
    I = circle_pos(p, round, nb);
    J = circle_pos(nb - 1 - p, round, nb);
    if (I > J) {
      const int t = I;
      I = J;
      J = t;
    }
  }
}

// One round with one workgroup per pair over all rows (the grouped batches
// and small single problems): the Gram and the rotation of X (and of V when
// accumulated) on FP64 MFMA, the inner solve as in svd_round_kernel.
// mode 0: full inner sweep on the fixed pairs; mode 1: cross-block inner steps.
__global__ __launch_bounds__(SR_THREADS, 4) void svd_round_mfma_kernel(SvdBatch Bt, int round, int mode) {
  __shared__ SvSmem sm;
  int prob = blockIdx.y, p = blockIdx.x;
  if (Bt.units) {
    const int2 u = Bt.units[blockIdx.x];
    prob = u.x;
    p = u.y;
  }
  const SvdProblem& P = Bt.arr ? Bt.arr[prob] : Bt.one;
  if (P.flag[1]) return;  // converged
  const int nb = P.nblk;
  if (mode == 1 && (nb <= 2 || round >= nb - 1)) return;
  if (p >= nb / 2) return;
  int I, J;
  round_pair(p, round, mode, nb, I, J);
  const int b = P.b, n2 = 2 * b;
  gram_mfma(sm, P.X, P.nr, 0, P.nr, I, J, b);
  const double abs_tol = fmax(4.930380657631324e-32 * (*P.fro2), 1e-300);
  if (mode == 1)
    jacobi_cross32(sm, P.tol, abs_tol);
  else
    jacobi_inner(sm, n2, P.tol, abs_tol);
  if (threadIdx.x == 0) {
    double mx = 0.0;
    for (int k = 0; k < n2 / 2; ++k) mx = fmax(mx, sm.rmax[k]);
    atomicMax(reinterpret_cast<unsigned long long*>(P.flag + 4), (unsigned long long)__double_as_longlong(mx));
  }
  if (!sm.any_outer) return;  // pair already orthogonal: W = I (uniform branch)
  if (threadIdx.x == 0) P.flag[0] = 1;
  apply_mfma(sm, P.X, P.nr, 0, P.nr, I, J, b);
  if (P.accum) apply_mfma(sm, P.V, P.ncp, 0, P.ncp, I, J, b);
}

// Rounds with each pair split over q workgroups by rows (phase 0: partial
// Grams; phase 1: reduce, inner solve, rotate the workgroup's row chunk), on
// the VALU with LDS staging.
// mode 0: full inner sweep on the fixed pairs (2p, 2p+1) (one such round per
// sweep); mode 1: cross-block inner steps on the circle pairs of `round`.
__global__ __launch_bounds__(SR_THREADS) void svd_round_kernel(SvdBatch Bt, int round, int phase, int mode) {
  __shared__ SvSmem sm;
  const SvdProblem& P = Bt.arr ? Bt.arr[blockIdx.y] : Bt.one;
  if (P.flag[1]) return;  // converged
  const int nb = P.nblk;
  // a single block pair (nb = 2) is swept whole in the mode-0 round; cross
  // rounds exist only for nb > 2, where every block is b = 16 wide
  if (mode == 1 && (nb <= 2 || round >= nb - 1)) return;
  const int q = P.q;
  const int item = blockIdx.x;
  const int p = item / q, c = item % q;
  if (p >= nb / 2) return;
  int I, J;
  round_pair(p, round, mode, nb, I, J);
  const int b = P.b, n2 = 2 * b;
  const int rcx = (P.nr + q - 1) / q;
  const int r0 = c * rcx, r1 = min(P.nr, r0 + rcx);
  double ar[GE_PER_THREAD], ai[GE_PER_THREAD];

  if (phase == 0) {
    gram_over_rows(sm, P.X, P.nr, r0, r1, I, J, b, ar, ai);
    double* out = P.Gpart + ((size_t)p * q + c) * (size_t)n2 * n2 * 2;
    if ((int)threadIdx.x < (n2 / 2) * (n2 / 2)) {
#pragma unroll
      for (int t = 0; t < GE_PER_THREAD; ++t) {
        int row, col;
        gram_slot(n2, t, row, col);
        const int e = col * n2 + row;
        out[2 * e] = ar[t];
        out[2 * e + 1] = ai[t];
      }
    }
    return;
  }
  {
    for (int t = 0; t < GE_PER_THREAD; ++t) {
      ar[t] = 0.0;
      ai[t] = 0.0;
    }
    // the q chunk partials in chunk order, four chunks' loads in flight at once
    if ((int)threadIdx.x < (n2 / 2) * (n2 / 2)) {
      int eo[GE_PER_THREAD];
#pragma unroll
      for (int t = 0; t < GE_PER_THREAD; ++t) {
        int row, col;
        gram_slot(n2, t, row, col);
        eo[t] = col * n2 + row;
      }
      const double2* base = reinterpret_cast<const double2*>(P.Gpart) + (size_t)p * q * n2 * n2;
      for (int cc = 0; cc < q; cc += 4) {
        double2 g[4][GE_PER_THREAD];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int t = 0; t < GE_PER_THREAD; ++t) {
            g[u][t].x = 0.0;
            g[u][t].y = 0.0;
            if (cc + u < q) g[u][t] = base[(size_t)(cc + u) * n2 * n2 + eo[t]];
          }
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (cc + u < q)
#pragma unroll
            for (int t = 0; t < GE_PER_THREAD; ++t) {
              ar[t] += g[u][t].x;
              ai[t] += g[u][t].y;
            }
      }
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < (n2 / 2) * (n2 / 2)) {
#pragma unroll
    for (int t = 0; t < GE_PER_THREAD; ++t) {
      int row, col;
      gram_slot(n2, t, row, col);
      sm.gr[row][col] = ar[t];
      sm.gi[row][col] = ai[t];
    }
  }
  __syncthreads();
  // One cyclic sweep of the 2b x 2b problem per pair visit (the outer sweeps
  // converge regardless; a full inner diagonalisation costs ~5x more).  Rotations are skipped only below eps^2 ||A||_F^2: a larger
  // absolute floor (e.g. eps ||A||_F^2) stops the late small-angle rotations
  // that give one-sided Jacobi its accuracy on small singular values.
  const double abs_tol = fmax(4.930380657631324e-32 * (*P.fro2), 1e-300);
  if (mode == 1)
    jacobi_cross32(sm, P.tol, abs_tol);
  else
    jacobi_inner(sm, n2, P.tol, abs_tol);
  // largest squared relative off-diagonal of this pair visit -> the sweep's maximum
  // (non-negative doubles order like their bit patterns)
  if (threadIdx.x == 0 && c == 0) {
    double mx = 0.0;
    for (int k = 0; k < n2 / 2; ++k) mx = fmax(mx, sm.rmax[k]);
    atomicMax(reinterpret_cast<unsigned long long*>(P.flag + 4), (unsigned long long)__double_as_longlong(mx));
  }
  if (!sm.any_outer) return;  // pair already orthogonal: W = I (uniform branch)
  if (threadIdx.x == 0 && c == 0) P.flag[0] = 1;
  apply_rows(sm, P.X, P.nr, r0, r1, I, J, b);
  if (!P.accum) return;
  const int rcv = (P.ncp + q - 1) / q;
  const int v0 = c * rcv;
  const int v1 = min(P.ncp, v0 + rcv);
  apply_rows(sm, P.V, P.ncp, v0, v1, I, J, b);
}

__global__ void svd_sweep_end_kernel(SvdBatch Bt, int64_t nprob) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nprob) return;
  const SvdProblem& P = Bt.arr ? Bt.arr[i] : Bt.one;
  if (P.flag[1]) return;
  // Converged when a sweep rotated nothing, or when its largest relative
  // off-diagonal was below sqrt(tol) / 2 (squared: tol / 4): one-sided
  // Jacobi converges quadratically, so after such a sweep every pair's
  // cosine is ~ that squared, below tol, and the verification sweep would
  // rotate nothing.
  unsigned long long* om = reinterpret_cast<unsigned long long*>(P.flag + 4);
  const double mx2 = __longlong_as_double((long long)*om);
  if (P.flag[0] == 0 || mx2 < 0.25 * P.tol) P.flag[1] = 1;
  P.flag[0] = 0;
  P.flag[2] += 1;
  *om = 0ull;
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
        if (P.pre) {  // X = R^H, R the upper triangle of W (ld m)
          if (c <= r) {
            const double2 a = reinterpret_cast<const double2*>(P.W)[(size_t)r * P.m + c];
            v.x = a.x;
            v.y = -a.y;
          }
        } else if (!P.transposed) {
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
    } else if (P.accum) {
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
    *reinterpret_cast<unsigned long long*>(P.flag + 4) = 0ull;
  }
}

// ||X||_F^2, one workgroup per problem, fixed summation order
// Column norms sqrt(sum_r |X[r][j]|^2) into sigma[j], one wave per column
// (grid.x covers ncp / 4 columns per problem).  Deterministic.
__global__ __launch_bounds__(SV_THREADS) void svd_colnorm_kernel(SvdBatch Bt) {
  const SvdProblem& P = Bt.arr ? Bt.arr[blockIdx.y] : Bt.one;
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * (SV_THREADS / 64) + (threadIdx.x >> 6);
  if (j >= P.ncp) return;
  const double2* col = reinterpret_cast<const double2*>(P.X) + (size_t)j * P.nr;
  double s = 0.0;
  for (int r = lane; r < P.nr; r += 64) {
    const double2 v = col[r];
    s += v.x * v.x + v.y * v.y;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) P.sigma[j] = sqrt(s);
}

// ||X||_F^2 from the column norms, one workgroup per problem, fixed order
__global__ __launch_bounds__(SV_THREADS) void svd_fro_kernel(SvdBatch Bt) {
  __shared__ double red[SV_THREADS / 64];
  const SvdProblem& P = Bt.arr ? Bt.arr[blockIdx.x] : Bt.one;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double s = 0.0;
  for (int j = tid; j < P.ncp; j += SV_THREADS) s += P.sigma[j] * P.sigma[j];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) red[wave] = s;
  __syncthreads();
  if (tid == 0) *P.fro2 = red[0] + red[1] + red[2] + red[3];
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
  // column norms were written to sigma by svd_colnorm_kernel
  for (int j = tid; j < np2; j += SV_THREADS) {
    key[j] = (j < ncp) ? P.sigma[j] : -1.0;
    idx[j] = j;
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
    if (P.S) {
      if (P.s_mode == 1) {
        reinterpret_cast<double2*>(P.S)[j] = make_double2(key[j], 0.0);
      } else if (P.s_mode == 2) {
        reinterpret_cast<float2*>(P.S)[j] = make_float2((float)key[j], 0.0f);
      } else {
        P.S[j] = key[j];
      }
    }
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

// QR preconditioning of the grouped problems with pre set: the column norms
// of A (one wave per column) into sigma (free until the Jacobi's own column
// norms), their descending order (ties by index) into cperm, and
// W[:, c] = A[:, cperm[c]] in complex128.  Sorting the columns first makes
// the unpivoted QR's R rows decrease in norm, as with column pivoting: on
// Vidal-gauge bond matrices the Jacobi on R^H then needs ~9 sweeps where the
// plain one needs ~14.
template <typename S2>
__global__ __launch_bounds__(SV_THREADS) void svd_pre_norm_kernel(SvdBatch Bt) {
  const SvdProblem& P = Bt.arr[blockIdx.y];
  if (!P.pre) return;
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * (SV_THREADS / 64) + (threadIdx.x >> 6);
  if (j >= P.n) return;
  const S2* col = static_cast<const S2*>(P.A) + (size_t)j * P.lda;
  double s = 0.0;
  for (int r = lane; r < P.m; r += 64) {
    const S2 v = col[r];
    s += (double)v.x * v.x + (double)v.y * v.y;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) P.sigma[j] = s;
}

__global__ __launch_bounds__(SV_THREADS) void svd_pre_sort_kernel(SvdBatch Bt) {
  __shared__ double key[SORT_MAX];
  __shared__ int idx[SORT_MAX];
  const SvdProblem& P = Bt.arr[blockIdx.x];
  if (!P.pre) return;
  const int tid = threadIdx.x, n = P.n;
  int np2 = 1;
  while (np2 < n) np2 <<= 1;
  for (int j = tid; j < np2; j += SV_THREADS) {
    key[j] = (j < n) ? P.sigma[j] : -1.0;
    idx[j] = j;
  }
  __syncthreads();
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
  for (int j = tid; j < n; j += SV_THREADS) P.cperm[j] = idx[j];
}

template <typename S2>
__global__ void svd_pre_pack_kernel(SvdBatch Bt) {
  const SvdProblem& P = Bt.arr[blockIdx.y];
  if (!P.pre) return;
  const size_t total = (size_t)P.m * P.n;
  const S2* A = static_cast<const S2*>(P.A);
  double2* W = reinterpret_cast<double2*>(P.W);
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const int r = (int)(e % P.m), c = (int)(e / P.m);
    const S2 a = A[(size_t)P.cperm[c] * P.lda + r];
    W[e] = make_double2((double)a.x, (double)a.y);
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
      if (!U) continue;  // U not requested
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
      // preconditioned: X's row c belongs to A's column cperm[c]
      Vh[(size_t)(P.pre ? P.cperm[cidx] : cidx) * P.ldvh + j] = o;
    }
  }
}

// Orthonormal completion of the factor built as X_j / sigma_j (U, or V^H's
// rows when the Jacobi ran on A^H): for kept singular values (j < k of the
// truncation) at or below 1e-8 sigma_0 that column carries X's rounding
// noise, not an orthonormal direction (exactly zero sigma: a zero column).  Each such column j, in
// order, is projected off columns 0..j-1 twice (classical Gram-Schmidt with
// reorthogonalisation) and normalised; if too little of it survives (below
// 1e-3 of its length before), the unit vector e_t (t = j, j+1, ... mod L)
// is used instead.  The result is an orthonormal completion as LAPACK
// returns it; the product U diag(S) V^H changes by at most
// 1e-8 sigma_0 * O(eps) per column.  Such columns form the tail [j0, k) (sigma
// descends); their projections off the head [0, j0) run in parallel
// (svd_complete_head_kernel), the serial part is the tail alone
// (svd_complete_kernel, one workgroup per problem).  A problem without such
// columns returns at once.
constexpr int CP_THREADS = 256;
constexpr int CP_MAX_K = 4096;
constexpr int CP_HEAD_BLOCKS = 32;  // workgroups per problem projecting tail columns off the head

template <typename S2>
struct Fac {
  S2* p;
  int64_t ld;
  bool adj;  // column j of F is conj(row j of V^H)
  __device__ __forceinline__ double2 get(int64_t r, int64_t j) const {
    const S2 v = adj ? p[r * ld + j] : p[j * ld + r];
    double2 o;
    o.x = (double)v.x;
    o.y = adj ? -(double)v.y : (double)v.y;
    return o;
  }
  __device__ __forceinline__ void put(int64_t r, int64_t j, double re, double im) const {
    S2 o;
    o.x = re;
    o.y = adj ? -im : im;
    if (adj)
      p[r * ld + j] = o;
    else
      p[j * ld + r] = o;
  }
};

__device__ __forceinline__ double cp_wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ double cp_block_sum(double v, double* red) {
  v = cp_wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < CP_THREADS / 64; ++w) t += red[w];
  return t;
}

// The factor to complete and its extent: U (m rows) or, when the Jacobi ran
// on A^H, the rows of V^H (n entries each, read as columns of F).
template <typename S2>
__device__ __forceinline__ bool cp_setup(const SvdProblem& P, Fac<S2>& F, int64_t& L, int& klim, int& j0) {
  const int kmax = P.ncols;
  if (kmax > CP_MAX_K) return false;
  if (!P.transposed) {
    if (!P.U) return false;
    F.p = static_cast<S2*>(P.U);
    F.ld = P.ldu;
    F.adj = false;
    L = P.m;
  } else {
    F.p = static_cast<S2*>(P.Vh);
    F.ld = P.ldvh;
    F.adj = true;
    L = P.n;
  }
  // only the kept columns (the truncation's k, written by the sort kernel)
  // are completed: the caller drops the others (a TEBD bond matrix of rank
  // 256 in 512 columns would otherwise orthogonalise 256 discarded ones)
  klim = P.k_dev ? (int)min((int64_t)kmax, *P.k_dev) : kmax;
  // sigma descends: the columns to complete are the tail [j0, klim)
  const double tau = P.sigma[0] * 1e-8;
  int lo = 0, hi = klim;
  while (lo < hi) {
    const int mid = (lo + hi) / 2;
    if (P.sigma[mid] > tau)
      lo = mid + 1;
    else
      hi = mid;
  }
  j0 = lo;
  return j0 < klim;
}

// Two passes (classical Gram-Schmidt with reorthogonalisation) of column j
// against columns [lo, hi) of F.
template <typename S2>
__device__ __forceinline__ void cp_project(const Fac<S2>& F, int64_t L, int lo, int hi, int j, double* cr, double* ci) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int pass = 0; pass < 2; ++pass) {
    // c = F[:, lo:hi]^H v
    if (!F.adj) {
      // columns contiguous: one wave per column, lanes over rows
      for (int i = lo + wave; i < hi; i += CP_THREADS / 64) {
        double sr = 0.0, si = 0.0;
        for (int64_t r = lane; r < L; r += 64) {
          const double2 a = F.get(r, i), v = F.get(r, j);
          sr += a.x * v.x + a.y * v.y;
          si += a.x * v.y - a.y * v.x;
        }
        sr = cp_wave_sum(sr);
        si = cp_wave_sum(si);
        if (lane == 0) {
          cr[i] = sr;
          ci[i] = si;
        }
      }
    } else {
      // rows of V^H: F's column index is the contiguous one, so one
      // thread per column i walks the rows (v[r] is the same for all)
      for (int i = lo + tid; i < hi; i += CP_THREADS) {
        double sr = 0.0, si = 0.0;
        for (int64_t r = 0; r < L; ++r) {
          const double2 a = F.get(r, i), v = F.get(r, j);
          sr += a.x * v.x + a.y * v.y;
          si += a.x * v.y - a.y * v.x;
        }
        cr[i] = sr;
        ci[i] = si;
      }
    }
    __syncthreads();
    // v -= F[:, lo:hi] c (one thread per row; with V^H's rows each thread
    // walks contiguous memory)
    for (int64_t r = tid; r < L; r += CP_THREADS) {
      double2 v = F.get(r, j);
      for (int i = lo; i < hi; ++i) {
        const double2 a = F.get(r, i);
        v.x -= a.x * cr[i] - a.y * ci[i];
        v.y -= a.x * ci[i] + a.y * cr[i];
      }
      F.put(r, j, v.x, v.y);
    }
    __syncthreads();
  }
}

// Step 1, in parallel over the tail columns (grid.x workgroups per problem):
// each column's length before any projection goes to n0 (the Jacobi work
// matrix X, free once V^H and U are written), then the column is projected
// off the head columns [0, j0), which are orthonormal to working accuracy.
template <typename S2>
__global__ __launch_bounds__(CP_THREADS) void svd_complete_head_kernel(SvdBatch Bt) {
  const SvdProblem& P = Bt.arr ? Bt.arr[blockIdx.y] : Bt.one;
  Fac<S2> F;
  int64_t L;
  int klim, j0;
  if (!cp_setup(P, F, L, klim, j0)) return;
  __shared__ double cr[CP_MAX_K], ci[CP_MAX_K];
  __shared__ double red[CP_THREADS / 64];
  double* n0s = P.X;
  for (int j = j0 + blockIdx.x; j < klim; j += gridDim.x) {
    double n0 = 0.0;
    for (int64_t r = threadIdx.x; r < L; r += CP_THREADS) {
      const double2 v = F.get(r, j);
      n0 += v.x * v.x + v.y * v.y;
    }
    n0 = sqrt(cp_block_sum(n0, red));
    if (threadIdx.x == 0) n0s[j] = n0;
    if (n0 > 0.0 && j0 > 0) cp_project(F, L, 0, j0, j, cr, ci);
    __syncthreads();
  }
}

// Step 2, one workgroup per problem: each tail column in order is projected
// off the tail columns before it and normalised; if too little of it
// survives (below 1e-3 of its length before step 1), the unit vector e_t
// (t = j, j+1, ... mod L) is projected off every column before it instead.
template <typename S2>
__global__ __launch_bounds__(CP_THREADS) void svd_complete_kernel(SvdBatch Bt) {
  const SvdProblem& P = Bt.arr ? Bt.arr[blockIdx.y] : Bt.one;
  Fac<S2> F;
  int64_t L;
  int klim, j0;
  if (!cp_setup(P, F, L, klim, j0)) return;
  __shared__ double cr[CP_MAX_K], ci[CP_MAX_K];
  __shared__ double red[CP_THREADS / 64];
  const int tid = threadIdx.x;
  const double* n0s = P.X;
  int64_t next_unit = 0;
  for (int j = j0; j < klim; ++j) {
    const double n0 = n0s[j];
    bool unit = !(n0 > 0.0);
    if (!unit) cp_project(F, L, j0, j, j, cr, ci);
    for (int attempt = 0; attempt < 16; ++attempt) {
      if (unit) {
        // e_t, t taken in turn from the unit vectors not yet tried
        const int64_t t = (j + next_unit++) % L;
        for (int64_t r = tid; r < L; r += CP_THREADS) F.put(r, j, r == t ? 1.0 : 0.0, 0.0);
        __syncthreads();
        cp_project(F, L, 0, j, j, cr, ci);
      }
      double nn = 0.0;
      for (int64_t r = tid; r < L; r += CP_THREADS) {
        const double2 v = F.get(r, j);
        nn += v.x * v.x + v.y * v.y;
      }
      nn = sqrt(cp_block_sum(nn, red));
      if (nn > 1e-3 * (unit ? 1.0 : n0)) {
        const double inv = 1.0 / nn;
        for (int64_t r = tid; r < L; r += CP_THREADS) {
          const double2 v = F.get(r, j);
          F.put(r, j, v.x * inv, v.y * inv);
        }
        __syncthreads();
        break;
      }
      unit = true;
    }
  }
}

// After a fixed number of captured sweeps: bit 0 of *status is set for a
// problem whose sweeps did not converge (the graph's caller reads it).
__global__ void svd_status_kernel(SvdBatch Bt, int32_t* status) {
  const SvdProblem& P = Bt.arr ? Bt.arr[blockIdx.x] : Bt.one;
  if (threadIdx.x == 0 && !P.flag[1]) atomicOr(status, 1);
}

#define TRY_HIP_COLNORM(b, nprob, max_ncp, s)                                                             \
  do {                                                                                                    \
    hipLaunchKernelGGL(svd_colnorm_kernel, dim3((unsigned)(((max_ncp) + SV_THREADS / 64 - 1) / (SV_THREADS / 64)), \
                                                (unsigned)(nprob)),                                       \
                       dim3(SV_THREADS), 0, s, b);                                                        \
    TENET_HIP_CHECK(hipGetLastError());                                                                   \
  } while (0)

int launch_svd_init(const SvdBatch& b, int64_t nprob, int64_t max_elems, int max_ncp, int dtype, hipStream_t s) {
  const unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>((max_elems + 255) / 256, 1024));
  if (dtype == TENET_C128)
    hipLaunchKernelGGL(svd_init_kernel<double2>, dim3(gx, (unsigned)nprob), dim3(256), 0, s, b);
  else
    hipLaunchKernelGGL(svd_init_kernel<float2>, dim3(gx, (unsigned)nprob), dim3(256), 0, s, b);
  TENET_HIP_CHECK(hipGetLastError());
  TRY_HIP_COLNORM(b, nprob, max_ncp, s);
  hipLaunchKernelGGL(svd_fro_kernel, dim3((unsigned)nprob), dim3(SV_THREADS), 0, s, b);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

int launch_svd_round(const SvdBatch& b, int64_t nprob, int max_items, int round, int phase, int mode,
                     hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
  if (b.units && phase != 2) {
    set_error("SVD round: unit lists need the fused round kernel");
    return TENET_ERR_UNSUPPORTED;
  }
  // with a unit list, max_items is its length
  const dim3 grid = b.units ? dim3((unsigned)max_items) : dim3((unsigned)max_items, (unsigned)nprob);
  if (phase == 2) {
    if (ev0)
      hipExtLaunchKernelGGL(svd_round_mfma_kernel, grid, dim3(SR_THREADS), 0, s, ev0, ev1, 0, b, round, mode);
    else
      hipLaunchKernelGGL(svd_round_mfma_kernel, grid, dim3(SR_THREADS), 0, s, b, round, mode);
  } else if (ev0) {
    hipExtLaunchKernelGGL(svd_round_kernel, grid, dim3(SR_THREADS), 0, s, ev0, ev1, 0, b, round, phase, mode);
  } else {
    hipLaunchKernelGGL(svd_round_kernel, grid, dim3(SR_THREADS), 0, s, b, round, phase, mode);
  }
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
  TRY_HIP_COLNORM(b, nprob, max_ncp, s);
  hipLaunchKernelGGL(svd_sort_kernel, dim3((unsigned)nprob), dim3(SV_THREADS), 0, s, b);
  TENET_HIP_CHECK(hipGetLastError());
  const unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>((max_elems + 255) / 256, 2048));
  if (dtype == TENET_C128)
    hipLaunchKernelGGL(svd_write_kernel<double2>, dim3(gx, (unsigned)nprob), dim3(256), 0, s, b);
  else
    hipLaunchKernelGGL(svd_write_kernel<float2>, dim3(gx, (unsigned)nprob), dim3(256), 0, s, b);
  TENET_HIP_CHECK(hipGetLastError());
  if (dtype == TENET_C128) {
    hipLaunchKernelGGL(svd_complete_head_kernel<double2>, dim3(CP_HEAD_BLOCKS, (unsigned)nprob), dim3(CP_THREADS), 0,
                       s, b);
    TENET_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(svd_complete_kernel<double2>, dim3(1, (unsigned)nprob), dim3(CP_THREADS), 0, s, b);
  } else {
    hipLaunchKernelGGL(svd_complete_head_kernel<float2>, dim3(CP_HEAD_BLOCKS, (unsigned)nprob), dim3(CP_THREADS), 0,
                       s, b);
    TENET_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(svd_complete_kernel<float2>, dim3(1, (unsigned)nprob), dim3(CP_THREADS), 0, s, b);
  }
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

int launch_svd_pre_pack(const SvdBatch& b, int64_t nprob, int max_n, int64_t max_elems, int dtype, hipStream_t s) {
  if (max_n > SORT_MAX) {
    set_error("SVD preconditioning: n exceeds the sort limit (4096)");
    return TENET_ERR_UNSUPPORTED;
  }
  const dim3 gn((unsigned)((max_n + SV_THREADS / 64 - 1) / (SV_THREADS / 64)), (unsigned)nprob);
  if (dtype == TENET_C128)
    hipLaunchKernelGGL(svd_pre_norm_kernel<double2>, gn, dim3(SV_THREADS), 0, s, b);
  else
    hipLaunchKernelGGL(svd_pre_norm_kernel<float2>, gn, dim3(SV_THREADS), 0, s, b);
  TENET_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(svd_pre_sort_kernel, dim3((unsigned)nprob), dim3(SV_THREADS), 0, s, b);
  TENET_HIP_CHECK(hipGetLastError());
  const unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>((max_elems + 255) / 256, 1024));
  if (dtype == TENET_C128)
    hipLaunchKernelGGL(svd_pre_pack_kernel<double2>, dim3(gx, (unsigned)nprob), dim3(256), 0, s, b);
  else
    hipLaunchKernelGGL(svd_pre_pack_kernel<float2>, dim3(gx, (unsigned)nprob), dim3(256), 0, s, b);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

// Single-matrix form of the column ordering (tenet_svd_thin's preconditioned
// path): T = A (adj = 0, rows x cols = m x n) or A^H (adj = 1, n x m);
// norms of T's columns, their descending order (ties by index) into perm, and
// Tp = T[:, perm] in complex128 (rows x cols, ld rows).
template <typename S2>
__global__ __launch_bounds__(SV_THREADS) void pre1_norm_kernel(const S2* __restrict__ A, int64_t lda, int rows,
                                                                int cols, int adj, double* __restrict__ nrm) {
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * (SV_THREADS / 64) + (threadIdx.x >> 6);
  if (j >= cols) return;
  double s = 0.0;
  for (int r = lane; r < rows; r += 64) {
    const S2 v = adj ? A[(size_t)r * lda + j] : A[(size_t)j * lda + r];
    s += (double)v.x * v.x + (double)v.y * v.y;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) nrm[j] = s;
}

__global__ __launch_bounds__(SV_THREADS) void pre1_sort_kernel(const double* __restrict__ nrm, int cols,
                                                                int32_t* __restrict__ perm) {
  __shared__ double key[SORT_MAX];
  __shared__ int idx[SORT_MAX];
  const int tid = threadIdx.x;
  int np2 = 1;
  while (np2 < cols) np2 <<= 1;
  for (int j = tid; j < np2; j += SV_THREADS) {
    key[j] = (j < cols) ? nrm[j] : -1.0;
    idx[j] = j;
  }
  __syncthreads();
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
  for (int j = tid; j < cols; j += SV_THREADS) perm[j] = idx[j];
}

template <typename S2>
__global__ void pre1_pack_kernel(const S2* __restrict__ A, int64_t lda, int rows, int cols, int adj,
                                 const int32_t* __restrict__ perm, double2* __restrict__ Tp) {
  const size_t total = (size_t)rows * cols;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const int r = (int)(e % rows), c = (int)(e / rows);
    const int pc = perm[c];
    const S2 v = adj ? A[(size_t)r * lda + pc] : A[(size_t)pc * lda + r];
    Tp[e] = make_double2((double)v.x, adj ? -(double)v.y : (double)v.y);
  }
}

// dst from src (complex128, rows x cols, ld lds) through the column order
// perm: mode 0 dst[:, perm[c]] = src[:, c] (ld ldd); mode 1 (conjugate
// transpose) dst[perm[c], r] = conj(src[r, c]).
__global__ void perm_scatter_kernel(const double2* __restrict__ src, int rows, int cols, int64_t lds,
                                    double2* __restrict__ dst, int64_t ldd, const int32_t* __restrict__ perm,
                                    int mode) {
  const size_t total = (size_t)rows * cols;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const int r = (int)(e % rows), c = (int)(e / rows);
    const double2 v = src[(size_t)c * lds + r];
    if (mode == 0)
      dst[(size_t)perm[c] * ldd + r] = v;
    else
      dst[(size_t)r * ldd + perm[c]] = make_double2(v.x, -v.y);
  }
}

int launch_pre1_order_pack(const void* A, int64_t lda, int rows, int cols, int adj, int dtype, double* nrm,
                           int32_t* perm, double* Tp, hipStream_t s) {
  if (cols > SORT_MAX) {
    set_error("SVD preconditioning: more than 4096 columns to order");
    return TENET_ERR_UNSUPPORTED;
  }
  const dim3 gn((unsigned)((cols + SV_THREADS / 64 - 1) / (SV_THREADS / 64)));
  if (dtype == TENET_C128)
    hipLaunchKernelGGL(pre1_norm_kernel<double2>, gn, dim3(SV_THREADS), 0, s, (const double2*)A, lda, rows, cols,
                       adj, nrm);
  else
    hipLaunchKernelGGL(pre1_norm_kernel<float2>, gn, dim3(SV_THREADS), 0, s, (const float2*)A, lda, rows, cols, adj,
                       nrm);
  TENET_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(pre1_sort_kernel, dim3(1), dim3(SV_THREADS), 0, s, nrm, cols, perm);
  TENET_HIP_CHECK(hipGetLastError());
  const size_t total = (size_t)rows * cols;
  const unsigned g = (unsigned)std::max<size_t>(1, std::min<size_t>((total + 255) / 256, 2048));
  if (dtype == TENET_C128)
    hipLaunchKernelGGL(pre1_pack_kernel<double2>, dim3(g), dim3(256), 0, s, (const double2*)A, lda, rows, cols, adj,
                       perm, (double2*)Tp);
  else
    hipLaunchKernelGGL(pre1_pack_kernel<float2>, dim3(g), dim3(256), 0, s, (const float2*)A, lda, rows, cols, adj,
                       perm, (double2*)Tp);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

int launch_perm_scatter(const double* src, int rows, int cols, int64_t lds, double* dst, int64_t ldd,
                        const int32_t* perm, int mode, hipStream_t s) {
  const size_t total = (size_t)rows * cols;
  if (total == 0) return TENET_OK;
  const unsigned g = (unsigned)std::max<size_t>(1, std::min<size_t>((total + 255) / 256, 2048));
  hipLaunchKernelGGL(perm_scatter_kernel, dim3(g), dim3(256), 0, s, (const double2*)src, rows, cols, lds,
                     (double2*)dst, ldd, perm, mode);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

int launch_svd_status(const SvdBatch& b, int64_t nprob, int32_t* status, hipStream_t s) {
  hipLaunchKernelGGL(svd_status_kernel, dim3((unsigned)nprob), dim3(64), 0, s, b, status);
  TENET_HIP_CHECK(hipGetLastError());
  return TENET_OK;
}

}  // namespace tenet
