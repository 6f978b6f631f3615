// Batched Hermitian eigensolver for the FP64 subspace problems (n <= 256),
// one workgroup per matrix, with a parallel tridiagonal stage:
//   1. Householder reduction A = Q T Q^H to a real tridiagonal T. The working
//      copy lives in the workspace; each column step reads the reduced row
//      k, builds the reflector, forms y = A v with every thread on a (column,
//      row-range) pair and applies the rank-2 update row by row with the
//      lanes along the row (coalesced). Reflector k is kept in row k.
//   2. Eigenvalues of T by quadrisection (three Sturm counts per step), one
//      lane per eigenvalue, on T scaled by a power of two to norm <= 1. The
//      Sturm count uses the division-free
//      three-term recurrence p_i = (d_i - x) p_{i-1} - e_{i-1}^2 p_{i-2}
//      (sign changes of p = eigenvalues below x), rescaled by 2^-600 or
//      2^600 every four steps, pivots below eps 2^-10 |p_{i-1}| replaced by
//      -eps 2^-10 p_{i-1} (a backward perturbation of d_i of that size).
//   3. Eigenvectors of T for the n_want lowest eigenvalues by inverse
//      iteration, one lane per eigenvector: three solves of (T - lam I) x = b
//      with the LU factorisation with partial pivoting of the tridiagonal
//      (zero pivots perturbed to eps), from a pseudo-random start vector that
//      differs per eigenvector; then modified Gram-Schmidt (twice) inside
//      clusters of eigenvalues closer than 1e-6 |T|, one wave per cluster.
//   4. Back-transformation V = Q X: each wave takes 4 eigenvector columns at
//      a time in registers (lanes along the rows) and applies the n - 1
//      reflectors with wave reductions only (no block barriers).
// Same input and output conventions as qmx_heev_small_grouped; E gets all n
// eigenvalues (ascending), V the first n_want eigenvectors.
#include <algorithm>
#include <mutex>

#include "qmx_common.h"

using namespace qmx;

namespace {

// Global-memory pointers carry address space 1, so that their accesses are
// global (not flat) loads and stores, counted apart from LDS accesses.
#define QMX_GLOBAL __attribute__((address_space(1)))
typedef double dv2 __attribute__((ext_vector_type(2)));
using gd2 = QMX_GLOBAL dv2;
using gd = QMX_GLOBAL double;

__device__ __forceinline__ double2 ld2(const gd2* p) {
  const dv2 v = *p;
  return double2{v.x, v.y};
}
__device__ __forceinline__ void st2(gd2* p, double2 a) {
  dv2 v;
  v.x = a.x;
  v.y = a.y;
  *p = v;
}

constexpr int kXThreads = 256;
constexpr int kXWaves = kXThreads / 64;
constexpr int kXMaxN = 256;
constexpr int kXIters = 2;      // inverse-iteration solves per eigenvector
constexpr int kBtCols = 4;                // back-transformation columns per wave pass
constexpr int kBtRows = kXMaxN / 64;      // rows per lane in the back-transformation
constexpr int kBtR = 36;                  // rows per wave, row-block back-transformation
constexpr int kStage = 8;                 // steps per register-staged block in phase 3
constexpr double kEps = 2.220446049250313e-16;
constexpr double kClusterRtol = 1e-6;  // final MGS inside clusters (relative to the block norm)
constexpr double kTightTol = 1e-12;    // numerically degenerate eigenvalues (same)

// Work area of one matrix: W (n x n complex: the working copy, then the
// reflectors in its rows), X, U1, U2, U3 (n x n_want real each, inverse
// iteration), d, e (n real) and tau (n complex) of the tridiagonal form, and
// a status word (nonzero: non-finite input).
__host__ __device__ inline int64_t x_ws_stride(int64_t n, int64_t n_want) {
  const int64_t bytes = n * n * 16 + 4 * n * n_want * 8 + n * 32 + 16;
  return (bytes + 255) / 256 * 256;
}

struct XWork {
  gd2* W;
  gd* X;
  gd* U1;
  gd* U2;
  gd* U3;
  gd* d;
  gd* e;
  gd2* tau;
  QMX_GLOBAL int* status;
};

__device__ __forceinline__ XWork x_work(unsigned char* work, int n, int nw) {
  XWork w;
  w.W = (gd2*)work;
  w.X = (gd*)(work + (int64_t)n * n * 16);
  w.U1 = w.X + (int64_t)n * nw;
  w.U2 = w.U1 + (int64_t)n * nw;
  w.U3 = w.U2 + (int64_t)n * nw;
  w.d = w.U3 + (int64_t)n * nw;
  w.e = w.d + n;
  w.tau = (gd2*)(w.e + n);
  w.status = (QMX_GLOBAL int*)(w.tau + n);
  return w;
}

// LDS of the reduction kernel: v, w (n complex), ypart (kXThreads complex),
// red (8 doubles), then the packed lower triangle of the trailing block.
// Reduction kernel threads: 512 for n_max > kRBigN (one matrix per CU when
// the packed block fills the LDS), 256 below.
constexpr int kRMaxThreads = 512;
constexpr int kRBigN = 96;
__host__ __device__ inline size_t r_lds_head(int64_t n) {
  return (size_t)(2 * n + kRMaxThreads) * 16 + 8 * 8;
}

constexpr size_t kLdsMax = 163840;

__device__ __forceinline__ double xwave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ double xblock_sum(double x, double* red) {
  x = xwave_sum(x);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wv] = x;
  __syncthreads();
  double s = 0;
#pragma unroll
  for (int i = 0; i < kXWaves; ++i) s += red[i];
  return s;
}

template <int kRWaves>
__device__ __forceinline__ double rblock_sum(double x, double* red) {
  x = xwave_sum(x);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wv] = x;
  __syncthreads();
  double s = 0;
#pragma unroll
  for (int i = 0; i < kRWaves; ++i) s += red[i];
  return s;
}

// Start vector of inverse iteration: a splitmix64 hash of (row, column)
// mapped to [-1, 1).
__device__ __forceinline__ double start_value(int i, int j) {
  uint64_t z = ((uint64_t)(uint32_t)j << 32 | (uint32_t)i) + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(int64_t)z * 0x1p-63;
}

// Numbers of eigenvalues of the scaled T (diagonal d, squared off-diagonal
// e2, e2[i] couples i and i + 1) below x[0], x[1] and x[2]: three
// independent recurrences interleaved, d and e2 read four rows at a time.
// Sturm counts of the scaled T at three points at once: the number of
// sign changes of p_0 = 1, p_i = (d_i - x) p_{i-1} - e2_{i-1} p_{i-2}
// (division free) is the number of eigenvalues below x. A zero p_{i-1}
// matters only where the next coupling e2_{i-1} is zero (the recurrence
// would stay at zero); there it is replaced by -piv p_{i-2}, a backward
// perturbation of d_{i-1}. Sign changes are read from the sign bits (a zero
// counts either way consistently); pairs (p_i, p_{i-1}) are renormalised
// through their exponent every four rows.
struct Sturm3 {
  double pm[3], pmm[3];
  int cnt[3];
  __device__ __forceinline__ void step(double di, double c2, bool split_next, const double* x) {
    constexpr double piv = kEps * 0x1p-10;
    double p[3];
#pragma unroll
    for (int t = 0; t < 3; ++t) p[t] = (di - x[t]) * pm[t] - c2 * pmm[t];
    if (__builtin_expect(split_next, false)) {
#pragma unroll
      for (int t = 0; t < 3; ++t) p[t] = p[t] == 0.0 ? -piv * pm[t] : p[t];
    }
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      const unsigned s = (unsigned)(__double2hiint(p[t]) ^ __double2hiint(pm[t]));
      cnt[t] += (int)(s >> 31);
      pmm[t] = pm[t];
      pm[t] = p[t];
    }
  }
  __device__ __forceinline__ void rescale() {
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      const int e = __builtin_amdgcn_frexp_exp(fmax(fabs(pm[t]), fabs(pmm[t])));  // 0 for 0
      pm[t] = __builtin_amdgcn_ldexp(pm[t], -e);
      pmm[t] = __builtin_amdgcn_ldexp(pmm[t], -e);
    }
  }
};

__device__ __forceinline__ void sturm_count3(const double* __restrict__ d,
                                             const double* __restrict__ e2, int n, const double* x,
                                             int* cnt) {
  Sturm3 st;
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    st.pm[t] = 1.0;
    st.pmm[t] = 0.0;
    st.cnt[t] = 0;
  }
  // e2[-1] = 0: row 0 has no coupling below; row n - 1 has no next row.
  int i = 0;
  for (; i + 4 <= n; i += 4) {
    double dd[4], cc[5];
#pragma unroll
    for (int q = 0; q < 4; ++q) dd[q] = d[i + q];
#pragma unroll
    for (int q = 0; q < 5; ++q) cc[q] = (i + q > 0 && i + q < n) ? e2[i + q - 1] : 1.0;
    cc[0] = i > 0 ? cc[0] : 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) st.step(dd[q], cc[q], cc[q + 1] == 0.0, x);
    st.rescale();
  }
  for (; i < n; ++i)
    st.step(d[i], i > 0 ? e2[i - 1] : 0.0, i + 1 < n && e2[i] == 0.0, x);
#pragma unroll
  for (int t = 0; t < 3; ++t) cnt[t] = st.cnt[t];
}

struct XLds {
  double2* v;
  double2* w;
  double2* tau;
  double2* ypart;  // [kXThreads]; the counts of the first bisection pass
  double* d;       // scaled diagonal after phase 2 starts
  double* e;
  double* e2;
  double* lam;     // scaled eigenvalues
  double* lsc;     // scaled norm of the diagonal block(s) each eigenvalue lies in
  double* red;     // [8]
  int* ctl;        // [4]
  int* pos;        // [n] position of each eigenvector in its tight group
};

__device__ __forceinline__ XLds x_lds(unsigned char* base, int n) {
  XLds L;
  L.v = reinterpret_cast<double2*>(base);
  L.w = L.v + n;
  L.tau = L.w + n;
  L.ypart = L.tau + n;
  L.d = reinterpret_cast<double*>(L.ypart + kXThreads);
  L.e = L.d + n;
  L.e2 = L.e + n;
  L.lam = L.e2 + n;
  L.lsc = L.lam + n;
  L.red = L.lsc + n;
  L.ctl = reinterpret_cast<int*>(L.red + 8);
  L.pos = L.ctl + 4;
  return L;
}

__host__ __device__ inline size_t x_lds_bytes(int64_t n) {
  return (size_t)(3 * n + kXThreads) * 16 + (size_t)(5 * n + 8) * 8 + (size_t)(4 + n) * sizeof(int);
}

// One inverse-iteration solve for eigenvalue lj of the scaled T (d, e in
// LDS): b = scale * Xc (the previous iterate), factorisation of T - lj I
// with partial pivoting carrying b along (y overwrites Xc, the factors go to
// u1 = 1 / pivot, u2, u3), then back substitution x_i = (y_i - u2_i x_{i+1} -
// u3_i x_{i+2}) u1_i into Xc. A pivot below pivmin (eps times the norm of
// the eigenvalue's block of T) is replaced by +-pivmin. Columns are strided
// by nw (one lane per column); loads and stores are staged kStage rows at a
// time. Returns 1 / |x|.
__device__ double inverse_solve(const XLds& L, int n, int nw, double lj, double pivmin,
                                double scale, gd* Xc, gd* u1, gd* u2, gd* u3) {
  double a = L.d[0] - lj, c = n > 1 ? L.e[0] : 0.0;
  double yk = Xc[0] * scale;
  for (int k0 = 0; k0 < n - 1; k0 += kStage) {
    double bb[kStage], o1[kStage], o2[kStage], o3[kStage], oy[kStage];
#pragma unroll
    for (int q = 0; q < kStage; ++q) {
      const int k = k0 + q;
      bb[q] = k + 1 < n ? Xc[(int64_t)(k + 1) * nw] * scale : 0.0;
    }
#pragma unroll
    for (int q = 0; q < kStage; ++q) {
      const int k = k0 + q;
      if (k < n - 1) {
        const double ek = L.e[k], dk1 = L.d[k + 1] - lj;
        const double ek1 = k + 2 < n ? L.e[k + 1] : 0.0;
        if (fabs(a) >= fabs(ek)) {
          if (fabs(a) < pivmin) a = a < 0.0 ? -pivmin : pivmin;
          const double ai = 1.0 / a, l = ek * ai;
          o1[q] = ai;
          o2[q] = c;
          o3[q] = 0.0;
          oy[q] = yk;
          yk = bb[q] - l * yk;
          a = dk1 - l * c;
          c = ek1;
        } else {
          const double l = a / ek;
          o1[q] = 1.0 / ek;
          o2[q] = dk1;
          o3[q] = ek1;
          oy[q] = bb[q];
          yk = yk - l * bb[q];
          a = c - l * dk1;
          c = -l * ek1;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < kStage; ++q) {
      const int k = k0 + q;
      if (k < n - 1) {
        u1[(int64_t)k * nw] = o1[q];
        u2[(int64_t)k * nw] = o2[q];
        u3[(int64_t)k * nw] = o3[q];
        Xc[(int64_t)k * nw] = oy[q];
      }
    }
  }
  if (fabs(a) < pivmin) a = a < 0.0 ? -pivmin : pivmin;
  u1[(int64_t)(n - 1) * nw] = 1.0 / a;
  u2[(int64_t)(n - 1) * nw] = 0.0;
  u3[(int64_t)(n - 1) * nw] = 0.0;
  Xc[(int64_t)(n - 1) * nw] = yk;
  double x1 = 0.0, x2 = 0.0, nrm = 0.0;
  for (int i0 = n - 1; i0 >= 0; i0 -= kStage) {
    double yv[kStage], f1[kStage], f2[kStage], f3[kStage];
#pragma unroll
    for (int q = 0; q < kStage; ++q) {
      const int i = i0 - q;
      if (i >= 0) {
        yv[q] = Xc[(int64_t)i * nw];
        f1[q] = u1[(int64_t)i * nw];
        f2[q] = u2[(int64_t)i * nw];
        f3[q] = u3[(int64_t)i * nw];
      }
    }
#pragma unroll
    for (int q = 0; q < kStage; ++q) {
      const int i = i0 - q;
      if (i >= 0) {
        const double xi = (yv[q] - f2[q] * x1 - f3[q] * x2) * f1[q];
        yv[q] = xi;
        nrm += xi * xi;
        x2 = x1;
        x1 = xi;
      }
    }
#pragma unroll
    for (int q = 0; q < kStage; ++q) {
      const int i = i0 - q;
      if (i >= 0) Xc[(int64_t)i * nw] = yv[q];
    }
  }
  return nrm > 0.0 && nrm <= 1.7e308 ? 1.0 / sqrt(nrm) : 1.0;
}

// Column j of X (n rows, row stride nw) normalised and orthogonalised
// against the orthonormal columns s..j-1 (classical Gram-Schmidt, twice),
// by one wave with the lanes along the rows.
__device__ void group_orthonormalise(gd* X, int n, int nw, int s, int j) {
  const int lane = threadIdx.x & 63;
  double x[kBtRows];
  double nn = 0;
#pragma unroll
  for (int r = 0; r < kBtRows; ++r) {
    const int i = lane + 64 * r;
    x[r] = i < n ? X[(int64_t)i * nw + j] : 0.0;
    nn += x[r] * x[r];
  }
  nn = xwave_sum(nn);
  double inv = nn > 0.0 ? 1.0 / sqrt(nn) : 0.0;
#pragma unroll
  for (int r = 0; r < kBtRows; ++r) x[r] *= inv;
  for (int pass = 0; pass < 2; ++pass) {
    for (int m = s; m < j; ++m) {
      double xm[kBtRows];
      double dt = 0;
#pragma unroll
      for (int r = 0; r < kBtRows; ++r) {
        const int i = lane + 64 * r;
        xm[r] = i < n ? X[(int64_t)i * nw + m] : 0.0;
        dt += x[r] * xm[r];
      }
      dt = xwave_sum(dt);
#pragma unroll
      for (int r = 0; r < kBtRows; ++r) x[r] -= dt * xm[r];
    }
    nn = 0;
#pragma unroll
    for (int r = 0; r < kBtRows; ++r) nn += x[r] * x[r];
    nn = xwave_sum(nn);
    inv = nn > 0.0 ? 1.0 / sqrt(nn) : 0.0;
#pragma unroll
    for (int r = 0; r < kBtRows; ++r) x[r] *= inv;
  }
#pragma unroll
  for (int r = 0; r < kBtRows; ++r) {
    const int i = lane + 64 * r;
    if (i < n) X[(int64_t)i * nw + j] = x[r];
  }
}

// Householder reduction of one matrix (phase 1). Steps k < ks run on the
// full Hermitian copy W in global memory; from step ks on, the lower
// triangle of the trailing block (rows and columns ks..n-1) lives packed in
// LDS (row a = i - ks at offset a (a + 1) / 2), and only the reflectors go
// back to global memory. ks is the first step whose trailing block fits in
// the `cap` complex elements of LDS left after the vectors.
template <int kRThreads>
__device__ void heevx_reduce_one(const double2* A_, int n, int64_t lda, unsigned char* work,
                                 int nw, int64_t cap) {
  extern __shared__ __align__(16) unsigned char smem_r[];
  struct {
    double2* v;
    double2* w;
    double2* ypart;
    double* red;
  } L;
  L.v = reinterpret_cast<double2*>(smem_r);
  L.w = L.v + n;
  L.ypart = L.w + n;
  L.red = reinterpret_cast<double*>(L.ypart + kRMaxThreads);
  double2* P = reinterpret_cast<double2*>(smem_r + r_lds_head(n));
  constexpr int kRWaves = kRThreads / 64;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const gd2* A = (const gd2*)A_;
  const XWork G = x_work(work, n, nw);
  gd2* W = G.W;
  int ms = n;
  while ((int64_t)ms * (ms + 1) / 2 > cap) --ms;
  const int ks = n - ms;  // steps 0..ks-1 in global memory

  // 0. Hermitian working copy from the lower triangle (global steps), or
  // the packed lower triangle straight into LDS.
  if (ks > 0) {
    for (int idx = tid; idx < n * n; idx += kRThreads) {
      const int i = idx / n, j = idx - i * n;
      double2 x;
      if (j < i) {
        x = ld2(A + (int64_t)i * lda + j);
      } else if (j > i) {
        const double2 t = ld2(A + (int64_t)j * lda + i);
        x = double2{t.x, -t.y};
      } else {
        x = double2{ld2(A + (int64_t)i * lda + i).x, 0.0};
      }
      st2(W + (int64_t)i * n + j, x);
    }
    __threadfence_block();
    __syncthreads();
  }

  // 1a. Householder steps on the global copy.
  for (int k = 0; k < ks; ++k) {
    const int m = n - k - 1;
    gd2* rowk = W + (int64_t)k * n;
    const double2 a1 = ld2(rowk + k + 1);  // W[k][k+1] = conj(alpha), read before it is overwritten
    double s = 0;
    for (int i = k + 2 + tid; i < n; i += kRThreads) {
      const double2 x = ld2(rowk + i);
      s += x.x * x.x + x.y * x.y;
    }
    const double sigma = rblock_sum<kRWaves>(s, L.red);
    const double2 alpha = double2{a1.x, -a1.y};
    double beta, sr = 0, si = 0;
    double2 tk;
    if (sigma == 0.0 && alpha.y == 0.0) {
      tk = double2{0.0, 0.0};
      beta = alpha.x;
    } else {
      const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + sigma);
      beta = alpha.x >= 0 ? -nrm : nrm;
      tk = double2{(beta - alpha.x) / beta, -alpha.y / beta};
      const double ar = alpha.x - beta, ai = alpha.y, den = ar * ar + ai * ai;
      sr = ar / den;
      si = -ai / den;
    }
    // v_i = x_i / (alpha - beta) with x_i = conj(W[k][i]); kept in LDS and in row k.
    for (int i = k + 2 + tid; i < n; i += kRThreads) {
      const double2 xr = ld2(rowk + i);
      const double xx = xr.x, xy = -xr.y;
      const double2 vv = double2{xx * sr - xy * si, xx * si + xy * sr};
      L.v[i] = vv;
      st2(rowk + i, vv);
    }
    if (tid == 0) {
      L.v[k + 1] = double2{1.0, 0.0};
      st2(rowk + k + 1, double2{1.0, 0.0});
      st2(G.tau + k, tk);
      G.e[k] = beta;
      G.d[k] = ld2(rowk + k).x;
    }
    __syncthreads();
    if (tk.x == 0.0 && tk.y == 0.0) continue;
    // y = A' v over the trailing block, A' Hermitian: y_i = sum_j conj(W[j][i]) v_j.
    int P = kRThreads / m;
    P = P > 8 ? 8 : (P < 1 ? 1 : P);
    const int chunk = (m + P - 1) / P;
    if (tid < P * m) {
      const int p = tid / m, i = k + 1 + (tid - p * m);
      const int j0 = k + 1 + p * chunk;
      const int j1 = min(n, j0 + chunk);
      double yr = 0, yi = 0;
      int j = j0;
      for (; j + 8 <= j1; j += 8) {
        double2 a[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) a[q] = ld2(W + (int64_t)(j + q) * n + i);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const double2 b = L.v[j + q];
          yr += a[q].x * b.x + a[q].y * b.y;
          yi += a[q].x * b.y - a[q].y * b.x;
        }
      }
      for (; j < j1; ++j) {
        const double2 a = ld2(W + (int64_t)j * n + i), b = L.v[j];
        yr += a.x * b.x + a.y * b.y;
        yi += a.x * b.y - a.y * b.x;
      }
      L.ypart[tid] = double2{yr, yi};
    }
    __syncthreads();
    double part = 0;
    for (int ii = tid; ii < m; ii += kRThreads) {
      double2 y = L.ypart[ii];
      for (int p = 1; p < P; ++p) {
        const double2 t = L.ypart[p * m + ii];
        y.x += t.x;
        y.y += t.y;
      }
      L.w[k + 1 + ii] = y;
      const double2 vi = L.v[k + 1 + ii];
      part += vi.x * y.x + vi.y * y.y;
    }
    const double vy = rblock_sum<kRWaves>(part, L.red);
    const double t2 = tk.x * tk.x + tk.y * tk.y;
    for (int i = k + 1 + tid; i < n; i += kRThreads) {
      const double2 y = L.w[i], vi = L.v[i];
      L.w[i] = double2{tk.x * y.x - tk.y * y.y - 0.5 * t2 * vy * vi.x,
                       tk.x * y.y + tk.y * y.x - 0.5 * t2 * vy * vi.y};
    }
    __syncthreads();
    // A' -= v w^H + w v^H: rows to waves, columns to lanes.
    {
      double2 vj[4], wj[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int j = k + 1 + lane + 64 * c;
        vj[c] = j < n ? L.v[j] : double2{0.0, 0.0};
        wj[c] = j < n ? L.w[j] : double2{0.0, 0.0};
      }
      for (int i0 = k + 1 + wave; i0 < n; i0 += 2 * kRWaves) {
        const int i1 = i0 + kRWaves;
        double2 a0[4], a1v[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int j = k + 1 + lane + 64 * c;
          if (j < n) {
            a0[c] = ld2(W + (int64_t)i0 * n + j);
            if (i1 < n) a1v[c] = ld2(W + (int64_t)i1 * n + j);
          }
        }
        {
          const double2 vi = L.v[i0], wi = L.w[i0];
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            const int j = k + 1 + lane + 64 * c;
            if (j < n) {
              double2 a = a0[c];
              a.x -= vi.x * wj[c].x + vi.y * wj[c].y + wi.x * vj[c].x + wi.y * vj[c].y;
              a.y -= vi.y * wj[c].x - vi.x * wj[c].y + wi.y * vj[c].x - wi.x * vj[c].y;
              st2(W + (int64_t)i0 * n + j, a);
            }
          }
        }
        if (i1 < n) {
          const double2 vi = L.v[i1], wi = L.w[i1];
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            const int j = k + 1 + lane + 64 * c;
            if (j < n) {
              double2 a = a1v[c];
              a.x -= vi.x * wj[c].x + vi.y * wj[c].y + wi.x * vj[c].x + wi.y * vj[c].y;
              a.y -= vi.y * wj[c].x - vi.x * wj[c].y + wi.y * vj[c].x - wi.x * vj[c].y;
              st2(W + (int64_t)i1 * n + j, a);
            }
          }
        }
      }
    }
    __threadfence_block();
    __syncthreads();
  }

  // 1b. The trailing block packed in LDS; steps ks..n-2 there.
  auto pk = [ks](int i, int j) { return (i - ks) * (i - ks + 1) / 2 + (j - ks); };
  for (int i = ks + wave; i < n; i += kRWaves) {
    for (int j = ks + lane; j <= i; j += 64) {
      double2 x;
      if (ks > 0) {
        x = ld2(W + (int64_t)i * n + j);
      } else {
        x = ld2(A + (int64_t)i * lda + j);
        if (j == i) x.y = 0.0;
      }
      P[pk(i, j)] = x;
    }
  }
  __syncthreads();
  for (int k = ks; k + 1 < n; ++k) {
    const int m = n - k - 1;
    gd2* rowk = W + (int64_t)k * n;
    const double2 alpha = P[pk(k + 1, k)];  // x_i = A[i][k], i > k
    // Every wave sums the column itself (the same lanes, the same order, so
    // every thread gets the same sigma and reflector without a block sum).
    double s = 0;
    for (int i = k + 2 + lane; i < n; i += 64) {
      const double2 x = P[pk(i, k)];
      s += x.x * x.x + x.y * x.y;
    }
    const double sigma = xwave_sum(s);
    double beta, sr = 0, si = 0;
    double2 tk;
    if (sigma == 0.0 && alpha.y == 0.0) {
      tk = double2{0.0, 0.0};
      beta = alpha.x;
    } else {
      const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + sigma);
      beta = alpha.x >= 0 ? -nrm : nrm;
      tk = double2{(beta - alpha.x) / beta, -alpha.y / beta};
      const double ar = alpha.x - beta, ai = alpha.y, den = ar * ar + ai * ai;
      sr = ar / den;
      si = -ai / den;
    }
    for (int i = k + 2 + tid; i < n; i += kRThreads) {
      const double2 x = P[pk(i, k)];
      const double2 vv = double2{x.x * sr - x.y * si, x.x * si + x.y * sr};
      L.v[i] = vv;
      st2(rowk + i, vv);
    }
    if (tid == 0) {
      L.v[k + 1] = double2{1.0, 0.0};
      st2(rowk + k + 1, double2{1.0, 0.0});
      st2(G.tau + k, tk);
      G.e[k] = beta;
      G.d[k] = P[pk(k, k)].x;
    }
    __syncthreads();
    if (tk.x == 0.0 && tk.y == 0.0) continue;
    // y = A' v with A' Hermitian from its lower triangle: A_ij = P(i, j)
    // for j <= i, conj(P(j, i)) for j > i.
    int Pp = kRThreads / m;
    Pp = Pp > 8 ? 8 : (Pp < 1 ? 1 : Pp);
    const int chunk = (m + Pp - 1) / Pp;
    if (tid < Pp * m) {
      const int p = tid / m, i = k + 1 + (tid - p * m);
      const int j0 = k + 1 + p * chunk;
      const int j1 = min(n, j0 + chunk);
      double yr = 0, yi = 0;
      const int jm = min(j1, i + 1);
      const int rb = (i - ks) * (i - ks + 1) / 2 - ks;  // row i: P[rb + j]
      double yr1 = 0, yi1 = 0;
      int j = j0;
      for (; j + 4 <= jm; j += 4) {  // row i, j <= i: A_ij = P(i, j)
        double2 a[4], b[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          a[q] = P[rb + j + q];
          b[q] = L.v[j + q];
        }
#pragma unroll
        for (int q = 0; q < 4; q += 2) {
          yr += a[q].x * b[q].x - a[q].y * b[q].y;
          yi += a[q].x * b[q].y + a[q].y * b[q].x;
          yr1 += a[q + 1].x * b[q + 1].x - a[q + 1].y * b[q + 1].y;
          yi1 += a[q + 1].x * b[q + 1].y + a[q + 1].y * b[q + 1].x;
        }
      }
      for (; j < jm; ++j) {
        const double2 a = P[rb + j], b = L.v[j];
        yr += a.x * b.x - a.y * b.y;
        yi += a.x * b.y + a.y * b.x;
      }
      // column i below the diagonal, j > i: A_ij = conj(P(j, i)); P(j, i) at
      // (j - ks)(j - ks + 1) / 2 + i - ks, the offset growing by j - ks + 1.
      j = max(j0, i + 1);
      int off = pk(j, i);
      for (; j + 4 <= j1; j += 4) {
        double2 a[4], b[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          a[q] = P[off];
          off += j + q - ks + 1;
          b[q] = L.v[j + q];
        }
#pragma unroll
        for (int q = 0; q < 4; q += 2) {
          yr += a[q].x * b[q].x + a[q].y * b[q].y;
          yi += a[q].x * b[q].y - a[q].y * b[q].x;
          yr1 += a[q + 1].x * b[q + 1].x + a[q + 1].y * b[q + 1].y;
          yi1 += a[q + 1].x * b[q + 1].y - a[q + 1].y * b[q + 1].x;
        }
      }
      for (; j < j1; ++j) {
        const double2 a = P[off], b = L.v[j];  // conj(a) * b
        off += j - ks + 1;
        yr += a.x * b.x + a.y * b.y;
        yi += a.x * b.y - a.y * b.x;
      }
      L.ypart[tid] = double2{yr + yr1, yi + yi1};
    }
    __syncthreads();
    // y_i = sum of the partial products; v^H y by every wave (same order on
    // every wave), then w = tau y - (|tau|^2 v^H y / 2) v.
    auto ysum = [&](int ii) {
      double2 y = L.ypart[ii];
      for (int p = 1; p < Pp; ++p) {
        const double2 t = L.ypart[p * m + ii];
        y.x += t.x;
        y.y += t.y;
      }
      return y;
    };
    double part = 0;
    for (int ii = lane; ii < m; ii += 64) {
      const double2 y = ysum(ii), vi = L.v[k + 1 + ii];
      part += vi.x * y.x + vi.y * y.y;
    }
    const double vy = xwave_sum(part);
    const double t2 = tk.x * tk.x + tk.y * tk.y;
    for (int ii = tid; ii < m; ii += kRThreads) {
      const double2 y = ysum(ii), vi = L.v[k + 1 + ii];
      L.w[k + 1 + ii] = double2{tk.x * y.x - tk.y * y.y - 0.5 * t2 * vy * vi.x,
                                tk.x * y.y + tk.y * y.x - 0.5 * t2 * vy * vi.y};
    }
    __syncthreads();
    // Lower triangle of A' -= v w^H + w v^H: rows to waves, columns to lanes.
    {
      // rows i = k + 1 + wave + 4 t, four at a time (all loads first), lanes
      // along the row j = k + 1 .. i (conflict-free).
      double2 vj[3], wj[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const int j = k + 1 + lane + 64 * c;
        vj[c] = j < n ? L.v[j] : double2{0.0, 0.0};
        wj[c] = j < n ? L.w[j] : double2{0.0, 0.0};
      }
      for (int i0 = k + 1 + wave; i0 < n; i0 += 4 * kRWaves) {
        double2 a[4][3];
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
          const int i = i0 + rr * kRWaves;
          const int rb = (i - ks) * (i - ks + 1) / 2 - ks;
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const int j = k + 1 + lane + 64 * c;
            a[rr][c] = (i < n && j <= i) ? P[rb + j] : double2{0.0, 0.0};
          }
        }
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
          const int i = i0 + rr * kRWaves;
          if (i >= n) break;
          const double2 vi = L.v[i], wi = L.w[i];
          const int rb = (i - ks) * (i - ks + 1) / 2 - ks;
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const int j = k + 1 + lane + 64 * c;
            if (j <= i) {
              double2 b = a[rr][c];
              b.x -= vi.x * wj[c].x + vi.y * wj[c].y + wi.x * vj[c].x + wi.y * vj[c].y;
              b.y -= vi.y * wj[c].x - vi.x * wj[c].y + wi.y * vj[c].x - wi.x * vj[c].y;
              P[rb + j] = b;
            }
          }
        }
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    G.d[n - 1] = P[pk(n - 1, n - 1)].x;
    G.e[n - 1] = 0.0;
    st2(G.tau + (n - 1), double2{0.0, 0.0});
  }
}

// Norm scale of eigenvalue x of the scaled T (d, e, e2 in LDS): T splits
// into diagonal blocks where a squared coupling is zero (the Sturm counts
// treat those rows as decoupled); the result is the largest Gershgorin norm
// among the blocks whose Gershgorin interval, widened by pad, holds x (1.0
// if none does).
__device__ double block_scale(const XLds& L, int n, double x, double pad) {
  double s = 0.0, blo = 0.0, bhi = 0.0, bn = 0.0;
  bool open = false;
  for (int i = 0; i < n; ++i) {
    const double r = (i > 0 ? fabs(L.e[i - 1]) : 0.0) + (i + 1 < n ? fabs(L.e[i]) : 0.0);
    const double lo = L.d[i] - r, hi = L.d[i] + r;
    const double m = fmax(fabs(lo), fabs(hi));
    blo = open ? fmin(blo, lo) : lo;
    bhi = open ? fmax(bhi, hi) : hi;
    bn = open ? fmax(bn, m) : m;
    open = true;
    if (i + 1 == n || L.e2[i] == 0.0) {  // the block ends at row i
      if (x >= blo - pad && x <= bhi + pad) s = fmax(s, bn);
      open = false;
    }
  }
  return s > 0.0 ? s : 1.0;
}

// Phases 2-4 of one matrix from the tridiagonal form in its work area.
__device__ void heevx_tail_one(int n, double* E_, double2* V_, int64_t ldv, unsigned char* work,
                               int nw, int* info) {
  extern __shared__ __align__(16) unsigned char smem_x[];
  const XLds L = x_lds(smem_x, n);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  gd* E = (gd*)E_;
  gd2* V = (gd2*)V_;
  const XWork G = x_work(work, n, nw);
  gd2* W = G.W;
  gd* X = G.X;
  gd* U1 = G.U1;
  gd* U2 = G.U2;
  gd* U3 = G.U3;
  for (int i = tid; i < n; i += kXThreads) {
    L.d[i] = G.d[i];
    L.e[i] = G.e[i];
    L.tau[i] = ld2(G.tau + i);
  }
  if (tid == 0) L.ctl[0] = 0;
  __syncthreads();

  // 2. Bisection on T scaled to |T| <= 1 (Gershgorin bound, power of two).
  double glo = 0.0, ghi = 0.0;
  bool finite = true;
  for (int i = 0; i < n; ++i) {
    const double r = (i > 0 ? fabs(L.e[i - 1]) : 0.0) + fabs(L.e[i]);
    const double lo = L.d[i] - r, hi = L.d[i] + r;
    finite = finite && fabs(lo) <= 1.7e308 && fabs(hi) <= 1.7e308;
    if (i == 0 || lo < glo) glo = lo;
    if (i == 0 || hi > ghi) ghi = hi;
  }
  const double tnorm = fmax(fabs(glo), fabs(ghi));
  if (!finite || !(tnorm <= 1.7e308)) {  // NaN or infinity in the input
    for (int j = tid; j < n; j += kXThreads) E[j] = __builtin_nan("");
    for (int idx = tid; idx < n * nw; idx += kXThreads) {
      const int i = idx / nw, j = idx - i * nw;
      st2(V + (int64_t)i * ldv + j, double2{__builtin_nan(""), 0.0});
    }
    if (tid == 0) {
      if (info) *info = 1;
      *G.status = 1;
    }
    return;
  }
  int ex = 0;
  if (tnorm > 0.0) frexp(tnorm, &ex);
  const double sc = ldexp(1.0, -ex), isc = ldexp(1.0, ex);  // sc * tnorm in [0.5, 1)
  __syncthreads();  // every thread has read the unscaled d and e
  for (int i = tid; i < n; i += kXThreads) {
    const double di = L.d[i] * sc, ei = L.e[i] * sc;
    L.d[i] = di;
    L.e[i] = ei;
    const double q = ei * ei;
    // Couplings below 2^-250 |T| split T (a relative perturbation below 1e-75);
    // then two rows of the Sturm recurrence shrink it by at most 2^-500 and
    // four rows (between renormalisations) stay above the underflow limit.
    L.e2[i] = q < 0x1p-500 ? 0.0 : q;
  }
  __syncthreads();
  // Each eigenvalue to 2 eps of its magnitude or of |T| (scaled to ~1),
  // then the norm of the diagonal block(s) of T it lies in: the scale of the
  // pivot floor of its inverse iteration and of the tight-group and cluster
  // tolerances below. Rows decoupled far above the wanted spectrum (the
  // dropped directions of a Davidson step, shifted by 1e6) then neither
  // perturb the wanted eigenvectors by eps |T| nor merge every wanted
  // eigenvalue into one cluster.
  {
    const double pad = 4.0 * kEps + 0x1p-1000;
    const double lo0 = glo * sc - pad, hi0 = ghi * sc + pad;
    auto bisect = [&](double& lo, double& hi, int j, double floor_tol) {
      for (int it = 0; it < 100; ++it) {
        const double tol = fmax(floor_tol, 2.0 * kEps * fmax(fabs(lo), fabs(hi)));
        const double h = 0.25 * (hi - lo);
        const double x[3] = {lo + h, lo + 2.0 * h, hi - h};
        if (hi - lo <= tol || !(x[0] > lo && x[2] < hi)) break;
        int c[3];
        sturm_count3(L.d, L.e2, n, x, c);
        if (c[0] > j) {
          hi = x[0];
        } else if (c[1] > j) {
          lo = x[0];
          hi = x[1];
        } else if (c[2] > j) {
          lo = x[1];
          hi = x[2];
        } else {
          lo = x[2];
        }
      }
    };
    // First pass, shared: the counts at 3 kXThreads evenly spaced points
    // (one Sturm pass of the workgroup; the reduction's ypart area is free
    // here). Eigenvalue j then starts from the points around it, an interval
    // 3 kXThreads + 1 times narrower; the search keeps count(lo) <= j <
    // count(hi), so the interval holds eigenvalue j.
    constexpr int kPts = 3 * kXThreads;
    int* c0 = reinterpret_cast<int*>(L.ypart);
    const double w0 = (hi0 - lo0) / (kPts + 1);
    {
      double x[3];
      int c[3];
#pragma unroll
      for (int t = 0; t < 3; ++t) x[t] = lo0 + (double)(tid + t * kXThreads + 1) * w0;
      sturm_count3(L.d, L.e2, n, x, c);
#pragma unroll
      for (int t = 0; t < 3; ++t) c0[tid + t * kXThreads] = c[t];
    }
    __syncthreads();
    for (int j = tid; j < n; j += kXThreads) {
      int a = -1, b = kPts;  // points a, b with count(a) <= j < count(b)
      while (b - a > 1) {
        const int m = (a + b) >> 1;
        if (c0[m] <= j) a = m; else b = m;
      }
      double lo = a < 0 ? lo0 : lo0 + (double)(a + 1) * w0;
      double hi = b >= kPts ? hi0 : lo0 + (double)(b + 1) * w0;
      bisect(lo, hi, j, 2.0 * kEps);
      const double s = block_scale(L, n, 0.5 * (lo + hi), 8.0 * kEps);
      const double lj = tnorm > 0.0 ? 0.5 * (lo + hi) : 0.0;
      L.lam[j] = lj;
      L.lsc[j] = s;
      E[j] = lj * isc;
    }
  }
  __syncthreads();

  // 3. Inverse iteration, one lane per wanted eigenvector. Eigenvalues
  // closer than kTightTol (numerically degenerate) form tight groups whose
  // members run in turn (round r = position in the group) and are
  // Gram-Schmidt orthogonalised against the earlier members after every
  // solve, so that each converges inside the orthogonal complement of the
  // others (well conditioned); the first members of all groups run together
  // in round 0 without barriers.
  for (int idx = tid; idx < n * nw; idx += kXThreads) {
    const int i = idx / nw, j = idx - i * nw;
    X[(int64_t)i * nw + j] = start_value(i, j);
  }
  if (tid == 0) L.ctl[1] = 0;
  __threadfence_block();
  __syncthreads();
  int mypos = 0;
  if (tid < nw) {
    while (tid - mypos > 0 &&
           L.lam[tid - mypos] - L.lam[tid - mypos - 1] <=
               fmax(kTightTol * fmax(L.lsc[tid - mypos], L.lsc[tid - mypos - 1]), 8.0 * kEps))
      ++mypos;
    L.pos[tid] = mypos;
    atomicMax(&L.ctl[1], mypos);
  }
  __syncthreads();
  const int rounds = L.ctl[1] + 1;
  for (int r = 0; r < rounds; ++r) {
    const bool active = tid < nw && mypos == r;
    double scale = 1.0;
    for (int it = 0; it < kXIters; ++it) {
      if (active)
        scale = inverse_solve(L, n, nw, L.lam[tid], fmax(kEps * L.lsc[tid], 0x1p-900), scale,
                              X + tid, U1 + tid, U2 + tid, U3 + tid);
      if (r > 0) {
        __threadfence_block();
        __syncthreads();
        int a = 0;
        for (int j = 0; j < nw; ++j) {
          if (L.pos[j] != r) continue;
          if ((a++ % kXWaves) == wave) group_orthonormalise(X, n, nw, j - r, j);
        }
        __threadfence_block();
        __syncthreads();
        scale = 1.0;
      }
    }
    if (r == 0 && active)
      for (int i = 0; i < n; ++i) X[(int64_t)i * nw + tid] *= scale;
  }
  __threadfence_block();
  __syncthreads();

  // Clusters of close eigenvalues: modified Gram-Schmidt, twice, one wave per
  // cluster, lanes along the rows.
  {
    int cl = 0;  // cluster counter (same on every lane)
    for (int s = 0; s < nw;) {
      int e = s + 1;
      while (e < nw && L.lam[e] - L.lam[e - 1] < kClusterRtol * fmax(L.lsc[e], L.lsc[e - 1])) ++e;
      if (e - s > 1 && (cl++ % kXWaves) == wave) {
        for (int pass = 0; pass < 2; ++pass) {
          for (int c = s; c < e; ++c) {
            double xc[kBtRows];
            double nn = 0;
#pragma unroll
            for (int r = 0; r < kBtRows; ++r) {
              const int i = lane + 64 * r;
              xc[r] = i < n ? X[(int64_t)i * nw + c] : 0.0;
              nn += xc[r] * xc[r];
            }
            nn = xwave_sum(nn);
            const double inv = nn > 0.0 ? 1.0 / sqrt(nn) : 0.0;
#pragma unroll
            for (int r = 0; r < kBtRows; ++r) {
              const int i = lane + 64 * r;
              xc[r] *= inv;
              if (i < n) X[(int64_t)i * nw + c] = xc[r];
            }
            for (int l = c + 1; l < e; ++l) {
              double xl[kBtRows];
              double dt = 0;
#pragma unroll
              for (int r = 0; r < kBtRows; ++r) {
                const int i = lane + 64 * r;
                xl[r] = i < n ? X[(int64_t)i * nw + l] : 0.0;
                dt += xc[r] * xl[r];
              }
              dt = xwave_sum(dt);
#pragma unroll
              for (int r = 0; r < kBtRows; ++r) {
                const int i = lane + 64 * r;
                if (i < n) X[(int64_t)i * nw + l] = xl[r] - dt * xc[r];
              }
            }
          }
        }
      }
      s = e;
    }
  }
  __threadfence_block();
  __syncthreads();

  if (tid == 0) {
    if (info) *info = 0;
    *G.status = 0;
  }
}

// Phase 4 of one matrix: V = Q X from the reflectors (rows of W), tau and the
// tridiagonal eigenvectors X in its work area (nothing when phase 2 found a
// non-finite input).
__device__ void heevx_back_one(int n, double2* V_, int64_t ldv, unsigned char* work, int nw) {
  extern __shared__ __align__(16) unsigned char smem_b[];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  gd2* V = (gd2*)V_;
  const XWork G = x_work(work, n, nw);
  if (*G.status != 0) return;
  gd2* W = G.W;
  gd* X = G.X;
  struct {
    double2* tau;
    double2* vb;
    double2* sp;
  } L;
  L.tau = reinterpret_cast<double2*>(smem_b);
  L.vb = L.tau + n;
  L.sp = L.vb + 3 * kXWaves * kBtR;
  for (int i = tid; i < n; i += kXThreads) L.tau[i] = ld2(G.tau + i);
  __syncthreads();
  // 4. V = Q X = H_0 H_1 ... H_{n-2} X.
  if (n <= kXWaves * kBtR) {
    // Waves own blocks of kBtR rows, lanes own columns (64 at a time): each
    // reflector is a partial dot per wave, one barrier, a sum over the waves
    // and the update. Reflectors go through a triple LDS buffer padded with
    // zeros to kXWaves * kBtR rows (v_{k-1} is stored while v_k is used,
    // v_{k-2} loaded into a register), the partial dots through a double one.
    constexpr int kPad = kXWaves * kBtR;
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    const int r0 = wv * kBtR;
    for (int g0 = 0; g0 < nw; g0 += 64) {
      const int j = g0 + lane;
      const bool jok = j < nw;
      double2 vr[kBtR];
#pragma unroll
      for (int r = 0; r < kBtR; ++r)
        vr[r] = double2{(r0 + r < n && jok) ? X[(int64_t)(r0 + r) * nw + j] : 0.0, 0.0};
      if (tid < kPad) {
        const int k = n - 2;
        L.vb[((k + 3) % 3) * kPad + tid] =
            (k >= 0 && tid > k && tid < n) ? ld2(W + (int64_t)k * n + tid) : double2{0.0, 0.0};
      }
      double2 vreg = (n >= 3 && tid < n && tid > n - 3) ? ld2(W + (int64_t)(n - 3) * n + tid)
                                                        : double2{0.0, 0.0};
      __syncthreads();
      for (int k = n - 2; k >= 0; --k) {
        const double2* vk = L.vb + (k % 3) * kPad + r0;
        if (k > 0 && tid < kPad) L.vb[((k - 1) % 3) * kPad + tid] = vreg;  // v_{k-1}
        vreg = (k > 1 && tid < n && tid > k - 2) ? ld2(W + (int64_t)(k - 2) * n + tid)
                                                 : double2{0.0, 0.0};  // v_{k-2}
        const bool live = r0 + kBtR - 1 > k;  // this wave has rows below k
        double pr = 0, pi = 0;
        if (live) {
#pragma unroll
          for (int r = 0; r < kBtR; ++r) {
            const double2 v = vk[r];
            pr += v.x * vr[r].x + v.y * vr[r].y;
            pi += v.x * vr[r].y - v.y * vr[r].x;
          }
        }
        L.sp[(k & 1) * (kXWaves * 64) + wv * 64 + lane] = double2{pr, pi};
        __syncthreads();
        double sr = 0, si = 0;
#pragma unroll
        for (int w = 0; w < kXWaves; ++w) {
          const double2 t = L.sp[(k & 1) * (kXWaves * 64) + w * 64 + lane];
          sr += t.x;
          si += t.y;
        }
        const double2 tk = L.tau[k];
        const double tr = tk.x * sr - tk.y * si, ti = tk.x * si + tk.y * sr;
        if (live) {
#pragma unroll
          for (int r = 0; r < kBtR; ++r) {
            const double2 v = vk[r];
            vr[r].x -= v.x * tr - v.y * ti;
            vr[r].y -= v.x * ti + v.y * tr;
          }
        }
      }
#pragma unroll
      for (int r = 0; r < kBtR; ++r)
        if (r0 + r < n && jok) st2(V + (int64_t)(r0 + r) * ldv + j, vr[r]);
      __syncthreads();
    }
    return;
  }
  // Larger n: 4 columns per wave pass, lanes along the rows, wave reductions.
  for (int c0 = kBtCols * wave; c0 < nw; c0 += kBtCols * kXWaves) {
    double2 vr[kBtRows][kBtCols];
#pragma unroll
    for (int r = 0; r < kBtRows; ++r) {
      const int i = lane + 64 * r;
#pragma unroll
      for (int q = 0; q < kBtCols; ++q) {
        const int j = c0 + q;
        vr[r][q] = double2{(i < n && j < nw) ? X[(int64_t)i * nw + j] : 0.0, 0.0};
      }
    }
    double2 vn[kBtRows];  // reflector k - 1, loaded one step ahead
#pragma unroll
    for (int r = 0; r < kBtRows; ++r) {
      const int i = lane + 64 * r, k = n - 2;
      vn[r] = (k >= 0 && i > k && i < n) ? ld2(W + (int64_t)k * n + i) : double2{0.0, 0.0};
    }
    for (int k = n - 2; k >= 0; --k) {
      double2 vk[kBtRows];
#pragma unroll
      for (int r = 0; r < kBtRows; ++r) {
        const int i = lane + 64 * r;
        vk[r] = vn[r];
        vn[r] = (k > 0 && i > k - 1 && i < n) ? ld2(W + (int64_t)(k - 1) * n + i) : double2{0.0, 0.0};
      }
      const double2 tk = L.tau[k];
      if (tk.x == 0.0 && tk.y == 0.0) continue;
      double sr[kBtCols], si[kBtCols];
#pragma unroll
      for (int q = 0; q < kBtCols; ++q) {
        double a = 0, b = 0;
#pragma unroll
        for (int r = 0; r < kBtRows; ++r) {
          a += vk[r].x * vr[r][q].x + vk[r].y * vr[r][q].y;
          b += vk[r].x * vr[r][q].y - vk[r].y * vr[r][q].x;
        }
        sr[q] = a;
        si[q] = b;
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int q = 0; q < kBtCols; ++q) {
          sr[q] += __shfl_xor(sr[q], off, 64);
          si[q] += __shfl_xor(si[q], off, 64);
        }
      }
#pragma unroll
      for (int q = 0; q < kBtCols; ++q) {
        const double tr = tk.x * sr[q] - tk.y * si[q], ti = tk.x * si[q] + tk.y * sr[q];
#pragma unroll
        for (int r = 0; r < kBtRows; ++r) {
          vr[r][q].x -= vk[r].x * tr - vk[r].y * ti;
          vr[r][q].y -= vk[r].x * ti + vk[r].y * tr;
        }
      }
    }
#pragma unroll
    for (int r = 0; r < kBtRows; ++r) {
      const int i = lane + 64 * r;
#pragma unroll
      for (int q = 0; q < kBtCols; ++q) {
        const int j = c0 + q;
        if (i < n && j < nw) st2(V + (int64_t)i * ldv + j, vr[r][q]);
      }
    }
  }
}

template <int NT>
__global__ void __launch_bounds__(NT)
    k_heevx_reduce(const qmx_heev_desc* __restrict__ desc, int64_t cap) {
  const qmx_heev_desc d = desc[blockIdx.x];
  const int nw = d.n_want > 0 && d.n_want < d.n ? d.n_want : d.n;
  heevx_reduce_one<NT>(static_cast<const double2*>(d.A), d.n, d.lda,
                       static_cast<unsigned char*>(d.work), nw, cap);
}

__global__ void __launch_bounds__(kXThreads) k_heevx_back(const qmx_heev_desc* __restrict__ desc) {
  const qmx_heev_desc d = desc[blockIdx.x];
  const int nw = d.n_want > 0 && d.n_want < d.n ? d.n_want : d.n;
  heevx_back_one(d.n, static_cast<double2*>(d.V), d.ldv, static_cast<unsigned char*>(d.work), nw);
}

__global__ void __launch_bounds__(kXThreads) k_heevx_tail(const qmx_heev_desc* __restrict__ desc) {
  const qmx_heev_desc d = desc[blockIdx.x];
  const int nw = d.n_want > 0 && d.n_want < d.n ? d.n_want : d.n;
  heevx_tail_one(d.n, d.E, static_cast<double2*>(d.V), d.ldv, static_cast<unsigned char*>(d.work),
                 nw, d.info);
}

}  // namespace

namespace {
// Dynamic LDS limits already set for this process (raised only when a larger
// launch needs it; the attribute call costs more than a launch).
size_t g_lds_reduce = 0, g_lds_reduce_big = 0, g_lds_tail = 0, g_lds_back = 0;
std::mutex g_lds_mu;

hipError_t set_lds_limit(const void* fn, size_t bytes, size_t* done) {
  std::lock_guard<std::mutex> lock(g_lds_mu);
  if (bytes <= *done) return hipSuccess;
  const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e == hipSuccess) *done = bytes;
  return e;
}
}  // namespace

extern "C" int64_t qmx_heevx_work_bytes(int64_t n, int64_t n_want) {
  if (n < 1 || n > kXMaxN) return -1;
  if (n_want <= 0 || n_want > n) n_want = n;
  return x_ws_stride(n, n_want);
}

extern "C" int qmx_heevx_grouped(const qmx_heev_desc* desc, int64_t count, int64_t n_max,
                                 void* stream) {
  QMX_CHECK_ARG(desc || count == 0, "qmx_heevx_grouped: null descriptors");
  QMX_CHECK_ARG(count >= 0 && count < (1ll << 31), "qmx_heevx_grouped: bad count");
  QMX_CHECK_ARG(n_max >= 1 && n_max <= kXMaxN, "qmx_heevx_grouped: n_max must be in 1..256");
  if (count == 0) return QMX_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  // Reduction: as much of the packed trailing block in LDS as fits.
  const size_t want = r_lds_head(n_max) + (size_t)16 * n_max * (n_max + 1) / 2;
  const size_t shr = want < kLdsMax ? want : kLdsMax;
  const int64_t cap = (int64_t)((shr - r_lds_head(n_max)) / 16);
  if (n_max > kRBigN) {
    QMX_HIP(set_lds_limit((const void*)k_heevx_reduce<kRMaxThreads>, shr, &g_lds_reduce_big));
    hipLaunchKernelGGL(k_heevx_reduce<kRMaxThreads>, dim3((unsigned)count), dim3(kRMaxThreads),
                       shr, st, desc, cap);
  } else {
    QMX_HIP(set_lds_limit((const void*)k_heevx_reduce<kXThreads>, shr, &g_lds_reduce));
    hipLaunchKernelGGL(k_heevx_reduce<kXThreads>, dim3((unsigned)count), dim3(kXThreads), shr,
                       st, desc, cap);
  }
  QMX_LAUNCHED();
  const size_t sh = x_lds_bytes(n_max);
  QMX_HIP(set_lds_limit((const void*)k_heevx_tail, sh, &g_lds_tail));
  hipLaunchKernelGGL(k_heevx_tail, dim3((unsigned)count), dim3(kXThreads), sh, st, desc);
  QMX_LAUNCHED();
  const size_t shb = (size_t)(n_max + 3 * kXWaves * kBtR + 2 * kXWaves * 64) * 16;
  QMX_HIP(set_lds_limit((const void*)k_heevx_back, shb, &g_lds_back));
  hipLaunchKernelGGL(k_heevx_back, dim3((unsigned)count), dim3(kXThreads), shb, st, desc);
  QMX_LAUNCHED();
  return QMX_OK;
}

// ---------------------------------------------------------------------------
// One larger Hermitian matrix (n <= 2048): the same tridiagonal method with
// the reduction spread over a grid of G co-resident workgroups.
//   1. k_ltrd (cooperative): rows of A are dealt cyclically (row i to
//      workgroup i mod G) and stay in LDS (in the work area when n is too
//      large for that). Column step k needs one grid-wide exchange: every
//      workgroup builds reflector k redundantly from column k of the
//      current matrix, owners form y_i = sum_j A_ij v_j and publish it with
//      A_i,k+1 and their part of v^H y; after the exchange every workgroup
//      forms w = tau y - |tau|^2 (v^H y) v / 2 and column k+1 of the updated
//      matrix (A_i,k+1 - v_i conj(w_k+1) - w_i), and owners apply the rank-2
//      update to their rows. The exchange buffers alternate between two
//      copies by step parity, and each workgroup signals a step by a flag of
//      its own (monotone step count), so the exchange needs no atomics; a
//      workgroup waits until every flag has reached the step. No dirty rows
//      stay in the L2 cache across the exchange when they live in LDS.
//   2. k_lbisect: all n eigenvalues by quadrisection, one lane each.
//   3. k_linvit: inverse iteration for the n_want lowest, one lane each.
//   4. k_lcluster: one wave per cluster of close eigenvalues: members of
//      numerically degenerate groups after the first are recomputed with
//      Gram-Schmidt against the earlier members after every solve, then the
//      cluster is orthonormalised by modified Gram-Schmidt, twice.
//   5. k_lback: V = Q X, reflectors staged through LDS for the whole
//      workgroup, eigenvector columns in registers (lanes along the rows).
namespace {

constexpr int kLThreads = 256;
constexpr int kLWaves = kLThreads / 64;
constexpr int kLMaxN = 2048;
constexpr int kLMaxG = 256;
constexpr unsigned kLSpinLimit = 1u << 22;  // polls (~1 s) before a wait gives up

// Offsets (bytes) of the arrays of the work area, computed once on the host.
struct LLayout {
  int64_t R, tau, d, e, ds, es, e2s, ybuf, cbuf, part, flags, lam, lsc, X, U1, U2, U3, rows;
};

struct LPtr {
  gd2* R;      // n x n: row k holds reflector k (entries k+1..n-1, v[k+1] = 1)
  gd2* tau;    // n
  gd* d;       // n, diagonal of T
  gd* e;       // n, off-diagonal of T (e[n-1] = 0)
  gd* ds;      // n, d and e scaled by a power of two to |T| < 1, and e^2
  gd* es;
  gd* e2s;
  gd2* ybuf;   // 2 x n
  gd2* cbuf;   // 2 x n
  gd* part;    // 2 x 4 kLMaxG (one per wave)
  QMX_GLOBAL unsigned* flags;  // kLMaxG step flags, [kLMaxG] abort, [kLMaxG + 1] status
  gd* lam;     // n scaled eigenvalues
  gd* lsc;     // n block norms
  gd* X;       // n x nw, U1..U3 likewise
  gd* U1;
  gd* U2;
  gd* U3;
  gd2* rows;   // G x nr x n rows (rows in the work area only)
};

__device__ __forceinline__ LPtr l_ptr(unsigned char* w, const LLayout& o) {
  LPtr p;
  p.R = (gd2*)(w + o.R);
  p.tau = (gd2*)(w + o.tau);
  p.d = (gd*)(w + o.d);
  p.e = (gd*)(w + o.e);
  p.ds = (gd*)(w + o.ds);
  p.es = (gd*)(w + o.es);
  p.e2s = (gd*)(w + o.e2s);
  p.ybuf = (gd2*)(w + o.ybuf);
  p.cbuf = (gd2*)(w + o.cbuf);
  p.part = (gd*)(w + o.part);
  p.flags = (QMX_GLOBAL unsigned*)(w + o.flags);
  p.lam = (gd*)(w + o.lam);
  p.lsc = (gd*)(w + o.lsc);
  p.X = (gd*)(w + o.X);
  p.U1 = (gd*)(w + o.U1);
  p.U2 = (gd*)(w + o.U2);
  p.U3 = (gd*)(w + o.U3);
  p.rows = (gd2*)(w + o.rows);
  return p;
}

__device__ __forceinline__ double wave_sum_fast(double v);

__device__ __forceinline__ double lblock_sum(double x, double* red) {
  x = wave_sum_fast(x);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wv] = x;
  __syncthreads();
  double s = 0;
#pragma unroll
  for (int i = 0; i < kLWaves; ++i) s += red[i];
  return s;
}

// The exchange between workgroups: data are written and read with
// agent-scope atomic accesses, which are coherent across the chip's L2
// caches by themselves (no cache write-back or invalidation, so rows kept in
// the work area stay cached); a wave waits for its data stores to complete
// before the workgroup signals the step with an agent-scope store of its
// flag, and readers load the data only after seeing every flag.
__device__ __forceinline__ void x_store(gd* p, double v) {
  __hip_atomic_store((double*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void x_store2(gd2* p, double2 v) {
  x_store((gd*)p, v.x);
  x_store((gd*)p + 1, v.y);
}
__device__ __forceinline__ double x_load(const gd* p) {
  return __hip_atomic_load((double*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double2 x_load2(const gd2* p) {
  return double2{x_load((const gd*)p), x_load((const gd*)p + 1)};
}

// Signal step `step` of this workgroup once the data stores of all its waves
// have completed (each wave waits for its vector memory operations, then the
// block barrier).
__device__ __forceinline__ void l_publish(QMX_GLOBAL unsigned* flags, unsigned step) {
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's stores are done
  __syncthreads();
  if (threadIdx.x == 0)
    __hip_atomic_store((unsigned*)(flags + blockIdx.x), step, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

// Wait until all G workgroups have signalled `step` (or the exchange was
// abandoned: a wait that exceeds kLSpinLimit polls sets the abort word, and
// every later wait returns at once, so the grid always drains). Wave 0 reads
// all flags at once each poll (at most kLMaxG / 64 per lane).
__device__ __forceinline__ void l_wait(QMX_GLOBAL unsigned* flags, int G, unsigned step) {
  if (threadIdx.x < 64) {
    unsigned* f = (unsigned*)flags;
    for (unsigned spin = 0;; ++spin) {
      unsigned lowest = 0xFFFFFFFFu;
#pragma unroll
      for (int t = 0; t < kLMaxG / 64; ++t) {
        const int q = threadIdx.x + 64 * t;
        const unsigned fv =
            q < G ? __hip_atomic_load(f + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0xFFFFFFFFu;
        lowest = fv < lowest ? fv : lowest;
      }
      const unsigned ab = __hip_atomic_load(f + kLMaxG, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (__all(lowest >= step) || ab != 0u) break;
      if (spin >= kLSpinLimit) {
        if (threadIdx.x == 0)
          __hip_atomic_store(f + kLMaxG, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  __syncthreads();
}

// Sum over the 64 lanes of a wave by DPP row rotations and row broadcasts
// (lane 63 ends with the total), returned to every lane.
__device__ __forceinline__ double dpp_add(double v, int ctrl_sel) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  int tl, th;
  switch (ctrl_sel) {
    case 0: tl = __builtin_amdgcn_update_dpp(0, lo, 0xB1, 0xF, 0xF, false);
            th = __builtin_amdgcn_update_dpp(0, hi, 0xB1, 0xF, 0xF, false); break;  // quad [1,0,3,2]
    case 1: tl = __builtin_amdgcn_update_dpp(0, lo, 0x4E, 0xF, 0xF, false);
            th = __builtin_amdgcn_update_dpp(0, hi, 0x4E, 0xF, 0xF, false); break;  // quad [2,3,0,1]
    case 2: tl = __builtin_amdgcn_update_dpp(0, lo, 0x124, 0xF, 0xF, false);
            th = __builtin_amdgcn_update_dpp(0, hi, 0x124, 0xF, 0xF, false); break;  // row_ror:4
    case 3: tl = __builtin_amdgcn_update_dpp(0, lo, 0x128, 0xF, 0xF, false);
            th = __builtin_amdgcn_update_dpp(0, hi, 0x128, 0xF, 0xF, false); break;  // row_ror:8
    case 4: tl = __builtin_amdgcn_update_dpp(0, lo, 0x142, 0xA, 0xF, false);
            th = __builtin_amdgcn_update_dpp(0, hi, 0x142, 0xA, 0xF, false); break;  // row_bcast:15
    default: tl = __builtin_amdgcn_update_dpp(0, lo, 0x143, 0xC, 0xF, false);
             th = __builtin_amdgcn_update_dpp(0, hi, 0x143, 0xC, 0xF, false); break;  // row_bcast:31
  }
  return v + __hiloint2double(th, tl);
}

__device__ __forceinline__ double wave_sum_fast(double v) {
#pragma unroll
  for (int c = 0; c < 6; ++c) v = dpp_add(v, c);
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
  return __hiloint2double(hi, lo);
}

__global__ void __launch_bounds__(64) k_lzero(unsigned char* work, LLayout lo) {
  QMX_GLOBAL unsigned* flags = l_ptr(work, lo).flags;
  for (int i = threadIdx.x; i < kLMaxG + 2; i += 64) flags[i] = 0u;
}

// Reflector of a column with leading entry alpha and sigma = |rest|^2 (as
// LAPACK's zlarfg): H = I - tau v v^H, v = (1, scale * rest), H^H x = beta e1.
struct LRefl {
  double2 tau;
  double beta, sre, sim;
};

__device__ __forceinline__ LRefl l_reflector(double2 alpha, double sigma) {
  LRefl r;
  r.sre = r.sim = 0.0;
  if (sigma == 0.0 && alpha.y == 0.0) {
    r.tau = double2{0.0, 0.0};
    r.beta = alpha.x;
  } else {
    const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + sigma);
    r.beta = alpha.x >= 0 ? -nrm : nrm;
    const double ib = 1.0 / r.beta;
    r.tau = double2{(r.beta - alpha.x) * ib, -alpha.y * ib};
    const double ar = alpha.x - r.beta, ai = alpha.y, iden = 1.0 / (ar * ar + ai * ai);
    r.sre = ar * iden;
    r.sim = -ai * iden;
  }
  return r;
}

// v (entries k+1..n-1) of reflector k from x into LDS; workgroup k mod G
// also stores it as row k of R with d[k], e[k] and tau[k].
__device__ __forceinline__ void l_store_reflector(const LPtr& P, const LRefl& r, const double2* x,
                                                  double2* v, int n, int k, int G, int g) {
  const bool mine = (k % G) == g;
  for (int j = k + 1 + (int)threadIdx.x; j < n; j += kLThreads) {
    double2 vj = double2{1.0, 0.0};
    if (j > k + 1) {
      const double2 t = x[j];
      vj = double2{t.x * r.sre - t.y * r.sim, t.x * r.sim + t.y * r.sre};
    }
    v[j] = vj;
    if (mine) st2(P.R + (int64_t)k * n + j, vj);
  }
  if (mine && threadIdx.x == 0) {
    P.d[k] = x[k].x;
    P.e[k] = r.beta;
    st2(P.tau + k, r.tau);
  }
}

constexpr int kLQ = kLMaxN / kLThreads;  // column entries per thread in the exchange

// Phase 1. LDS: v (2 x n complex, by step parity), w, x (n complex), red (16
// doubles), then nr rows of n complex when kRowsLds. Each step reads the
// exchange once (its parts, y and column k+1 together) and builds the next
// reflector before the rank-2 update, so the update and the next products
// need no further exchange.
template <bool kRowsLds>
__global__ void __launch_bounds__(kLThreads)
    k_ltrd(const double2* __restrict__ A_, int n, int64_t lda, unsigned char* work, LLayout lo,
           int nr) {
  extern __shared__ __align__(16) unsigned char smem_l[];
  const int G = gridDim.x, g = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const LPtr P = l_ptr(work, lo);
  const gd2* A = (const gd2*)A_;
  double2* vbuf = reinterpret_cast<double2*>(smem_l);
  double2* w = vbuf + 2 * n;
  double2* x = w + n;
  double* red = reinterpret_cast<double*>(x + n);
  double2* lrows = reinterpret_cast<double2*>(red + 16);
  gd2* grows = P.rows + (int64_t)g * nr * n;

  // Own rows, both triangles, from the lower triangle of A (UPLO = L).
  for (int r = 0; r < nr; ++r) {
    const int i = g + r * G;
    if (i >= n) break;
    for (int j = tid; j < n; j += kLThreads) {
      double2 a;
      if (j < i) {
        a = ld2(A + (int64_t)i * lda + j);
      } else if (j > i) {
        const double2 t = ld2(A + (int64_t)j * lda + i);
        a = double2{t.x, -t.y};
      } else {
        a = double2{ld2(A + (int64_t)i * lda + i).x, 0.0};
      }
      if constexpr (kRowsLds) lrows[r * n + j] = a;
      else st2(grows + (int64_t)r * n + j, a);
    }
  }
  double s0 = 0;
  for (int i = tid; i < n; i += kLThreads) {
    const double2 a = ld2(A + (int64_t)i * lda);
    x[i] = i == 0 ? double2{a.x, 0.0} : a;
    if (i >= 2) s0 += a.x * a.x + a.y * a.y;
  }
  double2 tau;  // of the current reflector, as built in every workgroup
  {
    const double sigma = lblock_sum(s0, red);
    const LRefl r0 = l_reflector(x[1], sigma);
    l_store_reflector(P, r0, x, vbuf, n, 0, G, g);
    tau = r0.tau;
  }
  __syncthreads();

  for (int k = 0; k < n - 1; ++k) {
    const int p = k & 1;
    const double2* v = vbuf + p * n;
    double2* vnext = vbuf + (1 - p) * n;
    // (b) Own rows i > k: y_i, the entry (i, k+1) and each wave's part of v^H y.
    double part = 0;
    for (int r = wave; r < nr; r += kLWaves) {
      const int i = g + r * G;
      if (i >= n) break;
      if (i <= k) continue;
      double yr = 0, yi = 0;
      if constexpr (kRowsLds) {
        const double2* Ri = lrows + r * n;
#pragma unroll 4
        for (int j = k + 1 + lane; j < n; j += 64) {
          const double2 a = Ri[j], b = v[j];
          yr += a.x * b.x - a.y * b.y;
          yi += a.x * b.y + a.y * b.x;
        }
      } else {
        const gd2* Ri = grows + (int64_t)r * n;
        for (int j = k + 1 + lane; j < n; j += 64) {
          const double2 a = ld2(Ri + j), b = v[j];
          yr += a.x * b.x - a.y * b.y;
          yi += a.x * b.y + a.y * b.x;
        }
      }
      yr = wave_sum_fast(yr);
      yi = wave_sum_fast(yi);
      if (lane == 0) {
        double2 c;
        if constexpr (kRowsLds) c = lrows[r * n + k + 1];
        else c = ld2(grows + (int64_t)r * n + k + 1);
        x_store2(P.ybuf + (int64_t)p * n + i, double2{yr, yi});
        x_store2(P.cbuf + (int64_t)p * n + i, c);
        const double2 vi = v[i];
        part += vi.x * yr + vi.y * yi;  // Re(conj(v_i) y_i)
      }
    }
    if (lane == 0) x_store(P.part + p * (4 * kLMaxG) + g * kLWaves + wave, part);
    l_publish(P.flags, (unsigned)(k + 1));
    l_wait(P.flags, G, (unsigned)(k + 1));

    // (d) One read of the exchange: the parts (wave 0), y and column k+1.
    double2 yq[kLQ], cq[kLQ];
#pragma unroll
    for (int q = 0; q < kLQ; ++q) {
      const int j = k + 1 + tid + q * kLThreads;
      if (j < n) {
        yq[q] = x_load2(P.ybuf + (int64_t)p * n + j);
        cq[q] = x_load2(P.cbuf + (int64_t)p * n + j);
      }
    }
    const double2 y1 = x_load2(P.ybuf + (int64_t)p * n + k + 1);
    if (tid < 64) {
      double pt[4 * kLMaxG / 64];
#pragma unroll
      for (int t = 0; t < 4 * kLMaxG / 64; ++t) {
        const int q = tid + 64 * t;
        pt[t] = q < G * kLWaves ? x_load(P.part + p * (4 * kLMaxG) + q) : 0.0;
      }
      double t = 0;
#pragma unroll
      for (int u = 0; u < 4 * kLMaxG / 64; ++u) t += pt[u];
      t = wave_sum_fast(t);
      if (tid == 0) red[12] = t;
    }
    __syncthreads();
    const double vy = red[12];
    const double t2 = tau.x * tau.x + tau.y * tau.y;
    const double2 w1 = double2{tau.x * y1.x - tau.y * y1.y - 0.5 * t2 * vy,
                               tau.x * y1.y + tau.y * y1.x};
    double sn = 0;  // sigma of the next reflector: |x_j|^2 over j >= k+3
#pragma unroll
    for (int q = 0; q < kLQ; ++q) {
      const int j = k + 1 + tid + q * kLThreads;
      if (j < n) {
        const double2 y = yq[q], c = cq[q], vj = v[j];
        const double2 wj = double2{tau.x * y.x - tau.y * y.y - 0.5 * t2 * vy * vj.x,
                                   tau.x * y.y + tau.y * y.x - 0.5 * t2 * vy * vj.y};
        w[j] = wj;
        // x_j = c_j - v_j conj(w_{k+1}) - w_j (v_{k+1} = 1)
        const double2 xj = double2{c.x - (vj.x * w1.x + vj.y * w1.y) - wj.x,
                                   c.y - (vj.y * w1.x - vj.x * w1.y) - wj.y};
        x[j] = xj;
        if (j >= k + 3) sn += xj.x * xj.x + xj.y * xj.y;
      }
    }
    const double sigma = lblock_sum(sn, red);  // also orders w and x before their readers
    double2 tau_next = double2{0.0, 0.0};
    if (k + 1 < n - 1) {
      const LRefl rn = l_reflector(x[k + 2], sigma);
      l_store_reflector(P, rn, x, vnext, n, k + 1, G, g);
      tau_next = rn.tau;
    }

    // (e) Rank-2 update of own rows i > k: A_ij -= v_i conj(w_j) + w_i conj(v_j).
    for (int r = wave; r < nr; r += kLWaves) {
      const int i = g + r * G;
      if (i >= n) break;
      if (i <= k) continue;
      const double2 vi = v[i], wi = w[i];
      if constexpr (kRowsLds) {
        double2* Ri = lrows + r * n;
#pragma unroll 4
        for (int j = k + 1 + lane; j < n; j += 64) {
          const double2 vj = v[j], wj = w[j];
          double2 a = Ri[j];
          a.x -= vi.x * wj.x + vi.y * wj.y + wi.x * vj.x + wi.y * vj.y;
          a.y -= vi.y * wj.x - vi.x * wj.y + wi.y * vj.x - wi.x * vj.y;
          Ri[j] = a;
        }
      } else {
        gd2* Ri = grows + (int64_t)r * n;
        for (int j = k + 1 + lane; j < n; j += 64) {
          const double2 vj = v[j], wj = w[j];
          double2 a = ld2(Ri + j);
          a.x -= vi.x * wj.x + vi.y * wj.y + wi.x * vj.x + wi.y * vj.y;
          a.y -= vi.y * wj.x - vi.x * wj.y + wi.y * vj.x - wi.x * vj.y;
          st2(Ri + j, a);
        }
      }
    }
    __syncthreads();  // the next reflector and the updated rows before step k + 1
    tau = tau_next;
  }
  if (g == 0 && tid == 0) {
    P.d[n - 1] = x[n - 1].x;
    P.e[n - 1] = 0.0;
  }
}

// Scaled T in LDS (d, e, e2) from the work area and the power-of-two scale;
// false for a non-finite T (or an abandoned reduction).
__device__ bool l_load_scaled(const LPtr& P, int n, double* d, double* e, double* e2, double* sc_out,
                              double* glo_out, double* ghi_out) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int i = tid; i < n; i += nt) {
    d[i] = P.d[i];
    e[i] = i + 1 < n ? P.e[i] : 0.0;
  }
  __syncthreads();
  double glo = 0.0, ghi = 0.0;
  bool finite = P.flags[kLMaxG] == 0u;
  for (int i = 0; i < n; ++i) {
    const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + fabs(e[i]);
    const double lo = d[i] - r, hi = d[i] + r;
    finite = finite && fabs(lo) <= 1.7e308 && fabs(hi) <= 1.7e308;
    if (i == 0 || lo < glo) glo = lo;
    if (i == 0 || hi > ghi) ghi = hi;
  }
  const double tnorm = fmax(fabs(glo), fabs(ghi));
  if (!finite || !(tnorm <= 1.7e308)) return false;
  int ex = 0;
  if (tnorm > 0.0) frexp(tnorm, &ex);
  const double sc = ldexp(1.0, -ex);
  __syncthreads();
  for (int i = tid; i < n; i += nt) {
    const double di = d[i] * sc, ei = e[i] * sc;
    d[i] = di;
    e[i] = ei;
    const double q = ei * ei;
    e2[i] = q < 0x1p-500 ? 0.0 : q;
  }
  __syncthreads();
  *sc_out = sc;
  *glo_out = glo;
  *ghi_out = ghi;
  return true;
}

// Phase 2: eigenvalue j = blockIdx.x * kLWaves + wave by multisection, one
// wave per eigenvalue: each round counts the eigenvalues below 192 evenly
// spaced points of the current interval (three per lane) and keeps the
// subinterval holding eigenvalue j, about seven rounds to 2 eps. Block 0
// also stores the scaled T for phases 3 and 4.
__global__ void __launch_bounds__(kLThreads)
    k_lbisect(int n, unsigned char* work, LLayout lo, double* __restrict__ E_, int* info) {
  extern __shared__ __align__(16) unsigned char smem_b2[];
  const LPtr P = l_ptr(work, lo);
  double* d = reinterpret_cast<double*>(smem_b2);
  double* e = d + n;
  double* e2 = e + n;
  const int tid = threadIdx.x, lane = tid & 63;
  const int j = blockIdx.x * kLWaves + (tid >> 6);
  gd* E = (gd*)E_;
  double sc, glo, ghi;
  if (!l_load_scaled(P, n, d, e, e2, &sc, &glo, &ghi)) {
    if (j < n && lane == 0) E[j] = __builtin_nan("");
    if (blockIdx.x == 0 && tid == 0) {
      P.flags[kLMaxG + 1] = 1u;
      if (info) *info = 1;
    }
    return;
  }
  if (blockIdx.x == 0) {
    for (int i = tid; i < n; i += kLThreads) {
      P.ds[i] = d[i];
      P.es[i] = e[i];
      P.e2s[i] = e2[i];
    }
    if (tid == 0) {
      P.flags[kLMaxG + 1] = 0u;
      if (info) *info = 0;
    }
  }
  if (j >= n) return;
  const double pad = 4.0 * kEps + 0x1p-1000;
  double lo_ = glo * sc - pad, hi_ = ghi * sc + pad;
  constexpr int kPts = 3 * 64;
  for (int round = 0; round < 40; ++round) {
    const double tol = fmax(2.0 * kEps, 2.0 * kEps * fmax(fabs(lo_), fabs(hi_)));
    const double h = (hi_ - lo_) / (kPts + 1);
    if (hi_ - lo_ <= tol || !(lo_ + h > lo_) || !(hi_ - h < hi_)) break;
    const double xq[3] = {lo_ + (double)(lane + 1) * h, lo_ + (double)(lane + 65) * h,
                          lo_ + (double)(lane + 129) * h};
    int c[3];
    sturm_count3(d, e2, n, xq, c);
    const int m = __popcll(__ballot(c[0] <= j)) + __popcll(__ballot(c[1] <= j)) +
                  __popcll(__ballot(c[2] <= j));
    const double nlo = m > 0 ? lo_ + (double)m * h : lo_;
    const double nhi = m < kPts ? lo_ + (double)(m + 1) * h : hi_;
    lo_ = nlo;
    hi_ = nhi;
  }
  if (lane == 0) {
    XLds L{};
    L.d = d;
    L.e = e;
    L.e2 = e2;
    const double lj = 0.5 * (lo_ + hi_);
    P.lam[j] = lj;
    P.lsc[j] = block_scale(L, n, lj, 8.0 * kEps);
    E[j] = lj / sc;
  }
}

__device__ __forceinline__ void l_load_t(const LPtr& P, int n, double* d, double* e) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    d[i] = P.ds[i];
    e[i] = P.es[i];
  }
  __syncthreads();
}

// One inverse-iteration solve of (T - lj I) x = scale * x for the scaled T
// (d, e) with all vectors in LDS (stride 1): LU with partial pivoting of the
// tridiagonal carrying the right side along (the factors to u1 = 1 / pivot,
// u2, u3, the eliminated right side over x), then back substitution into x.
// A pivot below pivmin is replaced by +-pivmin. Returns 1 / |x|.
__device__ double inverse_solve_lds(const double* d, const double* e, int n, double lj,
                                    double pivmin, double scale, double* x, double* u1, double* u2,
                                    double* u3) {
  double a = d[0] - lj, c = n > 1 ? e[0] : 0.0;
  double yk = x[0] * scale;
  for (int k = 0; k < n - 1; ++k) {
    const double bk = x[k + 1] * scale;
    const double ek = e[k], dk1 = d[k + 1] - lj;
    const double ek1 = k + 2 < n ? e[k + 1] : 0.0;
    if (fabs(a) >= fabs(ek)) {
      if (fabs(a) < pivmin) a = a < 0.0 ? -pivmin : pivmin;
      const double ai = 1.0 / a, l = ek * ai;
      u1[k] = ai;
      u2[k] = c;
      u3[k] = 0.0;
      x[k] = yk;
      yk = bk - l * yk;
      a = dk1 - l * c;
      c = ek1;
    } else {
      const double l = a / ek;
      u1[k] = 1.0 / ek;
      u2[k] = dk1;
      u3[k] = ek1;
      x[k] = bk;
      yk = yk - l * bk;
      a = c - l * dk1;
      c = -l * ek1;
    }
  }
  if (fabs(a) < pivmin) a = a < 0.0 ? -pivmin : pivmin;
  u1[n - 1] = 1.0 / a;
  u2[n - 1] = 0.0;
  u3[n - 1] = 0.0;
  x[n - 1] = yk;
  double x1 = 0.0, x2 = 0.0, nrm = 0.0;
  for (int i = n - 1; i >= 0; --i) {
    const double xi = (x[i] - u2[i] * x1 - u3[i] * x2) * u1[i];
    x[i] = xi;
    nrm += xi * xi;
    x2 = x1;
    x1 = xi;
  }
  return nrm > 0.0 && nrm <= 1.7e308 ? 1.0 / sqrt(nrm) : 1.0;
}

// Phase 3: eigenvectors blockIdx.x * ne + lane (lane < ne, j < nw), each with
// its iterate and factors in LDS (4n doubles after d and e), two solves from
// the start vector, normalised into column j of X.
__global__ void __launch_bounds__(64)
    k_linvit(int n, int nw, int ne, unsigned char* work, LLayout lo) {
  extern __shared__ __align__(16) unsigned char smem_i[];
  const LPtr P = l_ptr(work, lo);
  if (P.flags[kLMaxG + 1] != 0u) return;
  double* d = reinterpret_cast<double*>(smem_i);
  double* e = d + n;
  l_load_t(P, n, d, e);
  const int lane = threadIdx.x, j = blockIdx.x * ne + lane;
  if (lane >= ne || j >= nw) return;
  double* xv = e + n + (int64_t)lane * 4 * n;
  double* u1 = xv + n;
  double* u2 = u1 + n;
  double* u3 = u2 + n;
  for (int i = 0; i < n; ++i) xv[i] = start_value(i, j);
  const double pivmin = fmax(kEps * P.lsc[j], 0x1p-900), lj = P.lam[j];
  double scale = 1.0;
  for (int it = 0; it < kXIters; ++it)
    scale = inverse_solve_lds(d, e, n, lj, pivmin, scale, xv, u1, u2, u3);
  for (int i = 0; i < n; ++i) P.X[(int64_t)i * nw + j] = xv[i] * scale;
}

// Column c of X (stride nw) less its projections on columns s..c-1, then
// normalised; classical Gram-Schmidt (twice when `twice`), one wave.
__device__ void l_orthonormalise(gd* X, int n, int nw, int s, int c, bool twice) {
  const int lane = threadIdx.x & 63;
  for (int pass = 0; pass < (twice ? 2 : 1); ++pass) {
    for (int m = s; m < c; ++m) {
      double dt = 0;
      for (int i = lane; i < n; i += 64) dt += X[(int64_t)i * nw + c] * X[(int64_t)i * nw + m];
      dt = xwave_sum(dt);
      for (int i = lane; i < n; i += 64) X[(int64_t)i * nw + c] -= dt * X[(int64_t)i * nw + m];
    }
    double nn = 0;
    for (int i = lane; i < n; i += 64) nn += X[(int64_t)i * nw + c] * X[(int64_t)i * nw + c];
    nn = xwave_sum(nn);
    const double inv = nn > 0.0 ? 1.0 / sqrt(nn) : 0.0;
    for (int i = lane; i < n; i += 64) X[(int64_t)i * nw + c] *= inv;
  }
}

// Phase 4: wave q of the grid takes the cluster that starts at eigenvalue q.
__global__ void __launch_bounds__(kLThreads)
    k_lcluster(int n, int nw, unsigned char* work, LLayout lo) {
  extern __shared__ __align__(16) unsigned char smem_c[];
  const LPtr P = l_ptr(work, lo);
  if (P.flags[kLMaxG + 1] != 0u) return;
  double* d = reinterpret_cast<double*>(smem_c);
  double* e = d + n;
  l_load_t(P, n, d, e);
  const int lane = threadIdx.x & 63;
  const int s = blockIdx.x * kLWaves + (threadIdx.x >> 6);
  if (s >= nw) return;
  auto ctol = [&](int a, int b) { return kClusterRtol * fmax(P.lsc[a], P.lsc[b]); };
  auto tight = [&](int a, int b) {
    return P.lam[b] - P.lam[a] <= fmax(kTightTol * fmax(P.lsc[a], P.lsc[b]), 8.0 * kEps);
  };
  if (s > 0 && P.lam[s] - P.lam[s - 1] < ctol(s - 1, s)) return;  // not a cluster start
  int e_ = s + 1;
  while (e_ < nw && P.lam[e_] - P.lam[e_ - 1] < ctol(e_ - 1, e_)) ++e_;
  if (e_ - s < 2) return;
  XLds L{};
  L.d = d;
  L.e = e;
  // Degenerate groups: member c (not the first) restarts and is kept
  // orthogonal to the earlier members after each solve.
  int g0 = s;
  for (int c = s + 1; c < e_; ++c) {
    if (!tight(c - 1, c)) {
      g0 = c;
      continue;
    }
    for (int i = lane; i < n; i += 64) P.X[(int64_t)i * nw + c] = start_value(i, c);
    const double pivmin = fmax(kEps * P.lsc[c], 0x1p-900), lc = P.lam[c];
    for (int it = 0; it < kXIters; ++it) {
      if (lane == 0) inverse_solve(L, n, nw, lc, pivmin, 1.0, P.X + c, P.U1 + c, P.U2 + c, P.U3 + c);
      __threadfence_block();  // lane 0's stores before the other lanes read column c
      __builtin_amdgcn_wave_barrier();
      l_orthonormalise(P.X, n, nw, g0, c, false);
    }
  }
  // Modified Gram-Schmidt over the cluster, twice.
  for (int pass = 0; pass < 2; ++pass)
    for (int c = s; c < e_; ++c) l_orthonormalise(P.X, n, nw, s, c, false);
}

// Phase 5: V = H_0 H_1 ... H_{n-2} X for kC columns per wave. Reflectors
// (zero at and above their row) and their tau go through LDS kB at a time;
// the next batch is loaded into registers while the waves apply this one.
template <int kR, int kC, int kB>
__global__ void __launch_bounds__(kLThreads)
    k_lback(int n, int nw, unsigned char* work, LLayout lo, double2* __restrict__ V_, int64_t vsi,
            int64_t vsj) {
  extern __shared__ __align__(16) unsigned char smem_v[];
  const LPtr P = l_ptr(work, lo);
  double2* vb = reinterpret_cast<double2*>(smem_v);  // kB x n
  double2* tb = vb + kB * n;                          // kB taus
  gd2* V = (gd2*)V_;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int c0 = (blockIdx.x * kLWaves + wave) * kC;
  if (P.flags[kLMaxG + 1] != 0u) {  // non-finite input: NaN eigenvectors
    for (int r = 0; r < kR; ++r)
      for (int q = 0; q < kC; ++q) {
        const int i = lane + 64 * r, j = c0 + q;
        if (i < n && j < nw)
          st2(V + (int64_t)i * vsi + (int64_t)j * vsj, double2{__builtin_nan(""), 0.0});
      }
    return;
  }
  double2 xr[kR][kC];
#pragma unroll
  for (int r = 0; r < kR; ++r) {
    const int i = lane + 64 * r;
#pragma unroll
    for (int q = 0; q < kC; ++q) {
      const int j = c0 + q;
      xr[r][q] = double2{(i < n && j < nw) ? P.X[(int64_t)i * nw + j] : 0.0, 0.0};
    }
  }
  constexpr int kStg = (kR * 64 + kLThreads - 1) / kLThreads;  // staged entries per thread
  double2 stg[kB][kStg];
  double2 stau = double2{0.0, 0.0};
  // Batch t holds reflectors k = n - 2 - t kB - u, u = 0..kB-1 (k >= 0).
  auto fetch = [&](int t) {
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      const int k = n - 2 - t * kB - u;
#pragma unroll
      for (int q = 0; q < kStg; ++q) {
        const int i = tid + q * kLThreads;
        stg[u][q] = (k >= 0 && i < n && i > k) ? ld2(P.R + (int64_t)k * n + i) : double2{0.0, 0.0};
      }
    }
    const int kt = n - 2 - t * kB - tid;
    stau = (tid < kB && kt >= 0) ? ld2(P.tau + kt) : double2{0.0, 0.0};
  };
  auto put = [&]() {
#pragma unroll
    for (int u = 0; u < kB; ++u)
#pragma unroll
      for (int q = 0; q < kStg; ++q) {
        const int i = tid + q * kLThreads;
        if (i < n) vb[u * n + i] = stg[u][q];
      }
    if (tid < kB) tb[tid] = stau;
  };
  const int nbat = (n - 1 + kB - 1) / kB;
  fetch(0);
  put();
  __syncthreads();
  for (int t = 0; t < nbat; ++t) {
    if (t + 1 < nbat) fetch(t + 1);
#pragma unroll 1
    for (int u = 0; u < kB; ++u) {
      const int k = n - 2 - t * kB - u;
      if (k < 0) break;
      const double2 tk = tb[u];
      if (tk.x == 0.0 && tk.y == 0.0) continue;
      const double2* vk = vb + u * n;
      const int r0 = (k + 1) >> 6;  // rows below 64 r0 are zero in v_k
      double sr[kC], si[kC];
#pragma unroll
      for (int q = 0; q < kC; ++q) sr[q] = si[q] = 0.0;
#pragma unroll
      for (int r = 0; r < kR; ++r) {
        if (r < r0) continue;
        const int i = lane + 64 * r;
        const double2 vv = i < n ? vk[i] : double2{0.0, 0.0};
#pragma unroll
        for (int q = 0; q < kC; ++q) {
          sr[q] += vv.x * xr[r][q].x + vv.y * xr[r][q].y;
          si[q] += vv.x * xr[r][q].y - vv.y * xr[r][q].x;
        }
      }
#pragma unroll
      for (int c = 0; c < 6; ++c) {
#pragma unroll
        for (int q = 0; q < kC; ++q) {
          sr[q] = dpp_add(sr[q], c);
          si[q] = dpp_add(si[q], c);
        }
      }
#pragma unroll
      for (int q = 0; q < kC; ++q) {
        const double a = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(sr[q]), 63),
                                          __builtin_amdgcn_readlane(__double2loint(sr[q]), 63));
        const double b = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(si[q]), 63),
                                          __builtin_amdgcn_readlane(__double2loint(si[q]), 63));
        const double tr = tk.x * a - tk.y * b, ti = tk.x * b + tk.y * a;
#pragma unroll
        for (int r = 0; r < kR; ++r) {
          if (r < r0) continue;
          const int i = lane + 64 * r;
          const double2 vv = i < n ? vk[i] : double2{0.0, 0.0};
          xr[r][q].x -= vv.x * tr - vv.y * ti;
          xr[r][q].y -= vv.x * ti + vv.y * tr;
        }
      }
    }
    __syncthreads();
    if (t + 1 < nbat) {
      put();
      __syncthreads();
    }
  }
#pragma unroll
  for (int r = 0; r < kR; ++r) {
    const int i = lane + 64 * r;
#pragma unroll
    for (int q = 0; q < kC; ++q) {
      const int j = c0 + q;
      if (i < n && j < nw) st2(V + (int64_t)i * vsi + (int64_t)j * vsj, xr[r][q]);
    }
  }
}

}  // namespace

namespace {

LLayout l_layout(int64_t n, int64_t nw, int64_t rows_bytes, int64_t* total) {
  int64_t off = 0;
  auto take = [&](int64_t bytes) {
    const int64_t o = off;
    off += (bytes + 255) / 256 * 256;
    return o;
  };
  LLayout L;
  L.R = take(n * n * 16);
  L.tau = take(n * 16);
  L.d = take(n * 8);
  L.e = take(n * 8);
  L.ds = take(n * 8);
  L.es = take(n * 8);
  L.e2s = take(n * 8);
  L.ybuf = take(2 * n * 16);
  L.cbuf = take(2 * n * 16);
  L.part = take(2 * 4 * kLMaxG * 8);
  L.flags = take((kLMaxG + 2) * 4);
  L.lam = take(n * 8);
  L.lsc = take(n * 8);
  L.X = take(n * nw * 8);
  L.U1 = take(n * nw * 8);
  L.U2 = take(n * nw * 8);
  L.U3 = take(n * nw * 8);
  L.rows = take(rows_bytes);
  *total = off;
  return L;
}

size_t l_trd_lds(int64_t n, int64_t nr, bool rows_lds) {
  return (size_t)(4 * n) * 16 + 16 * 8 + (rows_lds ? (size_t)(nr * n) * 16 : 0);
}

// Rows stay in LDS when they fit with one row set per workgroup of the
// largest grid; else they live in the work area.
bool l_rows_fit(int64_t n) {
  const int64_t nr = (n + kLMaxG - 1) / kLMaxG;
  return l_trd_lds(n, nr, true) <= kLdsMax;
}

int64_t l_rows_reserve(int64_t n) { return l_rows_fit(n) ? 0 : (n + kLMaxG) * n * 16; }

size_t g_lds_ltrd = 0, g_lds_ltrd_g = 0, g_lds_lbis = 0, g_lds_linv = 0, g_lds_lclu = 0;
size_t g_lds_lback[8] = {0, 0, 0, 0, 0, 0, 0, 0};

template <int kR, int kC, int kB>
hipError_t launch_lback(int slot, int n, int nw, unsigned char* work, const LLayout& lo,
                        double2* V, int64_t vsi, int64_t vsj, hipStream_t st) {
  const size_t sh = (size_t)(kB * n + kB) * 16;
  hipError_t e = set_lds_limit((const void*)k_lback<kR, kC, kB>, sh, &g_lds_lback[slot]);
  if (e != hipSuccess) return e;
  const int grid = (nw + kLWaves * kC - 1) / (kLWaves * kC);
  hipLaunchKernelGGL((k_lback<kR, kC, kB>), dim3(grid), dim3(kLThreads), sh, st, n, nw, work, lo,
                     V, vsi, vsj);
  return hipGetLastError();
}

}  // namespace

extern "C" int64_t qmx_heevx_large_work_bytes(int64_t n, int64_t n_want) {
  if (n < 2 || n > kLMaxN) return -1;
  if (n_want <= 0 || n_want > n) n_want = n;
  int64_t total = 0;
  l_layout(n, n_want, l_rows_reserve(n), &total);
  return total;
}

extern "C" int qmx_heevx_large(const void* A, int64_t n, int64_t lda, int64_t n_want, double* E,
                               void* V, int64_t vsi, int64_t vsj, int* info, void* work,
                               int64_t work_bytes, int n_groups, void* stream) {
  QMX_CHECK_ARG(A && E && V && work, "qmx_heevx_large: null argument");
  QMX_CHECK_ARG(n >= 2 && n <= kLMaxN && lda >= n, "qmx_heevx_large: need 2 <= n <= 2048, lda >= n");
  if (n_want <= 0 || n_want > n) n_want = n;
  QMX_CHECK_ARG(work_bytes >= qmx_heevx_large_work_bytes(n, n_want),
                "qmx_heevx_large: work area too small");
  QMX_CHECK_ARG((reinterpret_cast<uintptr_t>(work) & 255) == 0,
                "qmx_heevx_large: work area must be 256-byte aligned");
  hipStream_t st = static_cast<hipStream_t>(stream);
  int dev = 0, n_cu = 0;
  QMX_HIP(hipGetDevice(&dev));
  QMX_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
  const int64_t gmax = std::min<int64_t>(kLMaxG, n_cu);
  int64_t G = n_groups > 0 ? n_groups : std::max<int64_t>(8, (n + 3) / 4);
  G = std::max<int64_t>(1, std::min<int64_t>({G, gmax, n}));
  int64_t nr = (n + G - 1) / G;
  bool rows_lds = l_trd_lds(n, nr, true) <= kLdsMax;
  if (!rows_lds) {  // more workgroups, fewer rows each, if that fits
    const int64_t nr2 = (n + gmax - 1) / gmax;
    if (l_trd_lds(n, nr2, true) <= kLdsMax) {
      nr = nr2;
      rows_lds = true;
    }
  }
  QMX_CHECK_ARG(rows_lds || l_rows_reserve(n) > 0,
                "qmx_heevx_large: rows need the work area on a device with fewer workgroups");
  G = (n + nr - 1) / nr;
  int64_t total = 0;
  const LLayout lo = l_layout(n, n_want, l_rows_reserve(n), &total);
  unsigned char* w = static_cast<unsigned char*>(work);
  const size_t sh = l_trd_lds(n, nr, rows_lds);
  const void* fn = rows_lds ? (const void*)k_ltrd<true> : (const void*)k_ltrd<false>;
  QMX_HIP(set_lds_limit(fn, sh, rows_lds ? &g_lds_ltrd : &g_lds_ltrd_g));
  int per_cu = 0;
  QMX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kLThreads, sh));
  QMX_CHECK_ARG((int64_t)per_cu * n_cu >= G, "qmx_heevx_large: grid cannot be co-resident");

  hipLaunchKernelGGL(k_lzero, dim3(1), dim3(64), 0, st, w, lo);
  QMX_LAUNCHED();
  const double2* Ap = static_cast<const double2*>(A);
  int ni = (int)n, nri = (int)nr;
  LLayout lo_arg = lo;
  void* args[] = {(void*)&Ap, (void*)&ni, (void*)&lda, (void*)&w, (void*)&lo_arg, (void*)&nri};
  QMX_HIP(hipLaunchCooperativeKernel(fn, dim3((unsigned)G), dim3(kLThreads), args, (unsigned)sh, st));

  const int nw = (int)n_want;
  QMX_HIP(set_lds_limit((const void*)k_lbisect, (size_t)n * 24, &g_lds_lbis));
  hipLaunchKernelGGL(k_lbisect, dim3((unsigned)((n + kLWaves - 1) / kLWaves)), dim3(kLThreads),
                     (size_t)n * 24, st, ni, w, lo, E, info);
  QMX_LAUNCHED();
  // Eigenvectors per inverse-iteration workgroup: their LDS (4n doubles
  // each, after d and e) within half the CU's LDS, at most 16.
  const int64_t ne = std::max<int64_t>(1, std::min<int64_t>(16, ((int64_t)kLdsMax / 2 - 16 * n) / (32 * n)));
  const size_t shi = (size_t)(2 * n + 4 * n * ne) * 8;
  QMX_HIP(set_lds_limit((const void*)k_linvit, shi, &g_lds_linv));
  hipLaunchKernelGGL(k_linvit, dim3((unsigned)((nw + ne - 1) / ne)), dim3(64), shi, st, ni, nw,
                     (int)ne, w, lo);
  QMX_LAUNCHED();
  QMX_HIP(set_lds_limit((const void*)k_lcluster, (size_t)n * 16, &g_lds_lclu));
  hipLaunchKernelGGL(k_lcluster, dim3((unsigned)((nw + kLWaves - 1) / kLWaves)), dim3(kLThreads),
                     (size_t)n * 16, st, ni, nw, w, lo);
  QMX_LAUNCHED();
  double2* Vp = static_cast<double2*>(V);
  const int64_t rows64 = (n + 63) / 64;
  hipError_t e;
  if (rows64 <= 6) e = launch_lback<6, 4, 8>(0, ni, nw, w, lo, Vp, vsi, vsj, st);
  else if (rows64 <= 8) e = launch_lback<8, 3, 8>(1, ni, nw, w, lo, Vp, vsi, vsj, st);
  else if (rows64 <= 10) e = launch_lback<10, 2, 8>(2, ni, nw, w, lo, Vp, vsi, vsj, st);
  else if (rows64 <= 12) e = launch_lback<12, 2, 8>(3, ni, nw, w, lo, Vp, vsi, vsj, st);
  else if (rows64 <= 16) e = launch_lback<16, 2, 4>(4, ni, nw, w, lo, Vp, vsi, vsj, st);
  else if (rows64 <= 20) e = launch_lback<20, 1, 4>(5, ni, nw, w, lo, Vp, vsi, vsj, st);
  else if (rows64 <= 24) e = launch_lback<24, 1, 4>(6, ni, nw, w, lo, Vp, vsi, vsj, st);
  else e = launch_lback<32, 1, 2>(7, ni, nw, w, lo, Vp, vsi, vsj, st);
  QMX_HIP(e);
  return QMX_OK;
}

// ---------------------------------------------------------------------------
// Cholesky factor and its inverse for one Hermitian positive definite matrix
// O (n <= 2048), in one cooperative launch over G workgroups that own rows
// of O cyclically (rows in the work area), exchanging one panel of nb columns
// at a time as in k_ltrd:
//   phase 1 (O = L L^H, right-looking): owners publish the panel's columns
//     of their rows; every workgroup factors the nb x nb diagonal block and
//     forms the panel's rows of L (L_j = O_j L_kk^-H); owners store their
//     rows' panel entries and update their trailing entries
//     O_ij -= sum_c L_ic conj(L_jc), j <= i;
//   phase 2 (X = L^-1, block forward substitution on R = I): owners of the
//     block's rows publish those rows of R and of the diagonal block of L;
//     every workgroup forms X_b = L_bb^-1 R_b, workgroup b mod G writes those
//     rows of X, and owners update their later rows R_i -= L_i,b X_b.
// info = 0, or k + 1 for the first non-positive (or non-finite) pivot k, as
// LAPACK's potrf; X is then not meaningful.
namespace {

struct CLayout {
  int64_t Lrows, Rrows, xbuf, flags;
};

__device__ __forceinline__ void c_publish(QMX_GLOBAL unsigned* flags, unsigned step) {
  l_publish(flags, step);
}

template <int kNb>
__global__ void __launch_bounds__(kLThreads)
    k_lpotrf(const double2* __restrict__ O_, int n, int64_t ldo, double2* __restrict__ X_,
             unsigned char* work, CLayout lo, int nr, int* info) {
  extern __shared__ __align__(16) unsigned char smem_p[];
  const int G = gridDim.x, g = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const gd2* O = (const gd2*)O_;
  gd2* Xo = (gd2*)X_;
  gd2* Lr = (gd2*)(work + lo.Lrows) + (int64_t)g * nr * n;  // own rows of O, then of L
  gd2* Rr = (gd2*)(work + lo.Rrows) + (int64_t)g * nr * n;  // own rows of R
  gd2* xb = (gd2*)(work + lo.xbuf);                          // 2 x (n kNb + kNb^2)
  QMX_GLOBAL unsigned* flags = (QMX_GLOBAL unsigned*)(work + lo.flags);
  const int64_t xstride = (int64_t)n * kNb + kNb * kNb;
  double2* pan = reinterpret_cast<double2*>(smem_p);  // n x kNb (phase 1) / kNb x n (phase 2)
  double2* dg = pan + (int64_t)n * kNb;               // kNb x kNb
  int* fail = reinterpret_cast<int*>(dg + kNb * kNb);
  if (tid == 0) *fail = 0;

  for (int r = 0; r < nr; ++r) {
    const int i = g + r * G;
    if (i >= n) break;
    for (int j = tid; j < n; j += kLThreads) {
      st2(Lr + (int64_t)r * n + j, j <= i ? ld2(O + (int64_t)i * ldo + j) : double2{0.0, 0.0});
      st2(Rr + (int64_t)r * n + j, double2{j == i ? 1.0 : 0.0, 0.0});
    }
  }
  __syncthreads();
  const int nbk = (n + kNb - 1) / kNb;
  unsigned step = 0;

  // Phase 1.
  for (int b = 0; b < nbk; ++b) {
    const int c0 = b * kNb, w = min(kNb, n - c0);
    gd2* buf = xb + (int64_t)(step & 1) * xstride;
    for (int r = wave; r < nr; r += kLWaves) {
      const int i = g + r * G;
      if (i >= n) break;
      if (i < c0) continue;
      for (int c = lane; c < w; c += 64)
        x_store2(buf + (int64_t)(i - c0) * kNb + c,
                 c0 + c <= i ? ld2(Lr + (int64_t)r * n + c0 + c) : double2{0.0, 0.0});
    }
    ++step;
    l_publish(flags, step);
    l_wait(flags, G, step);
    for (int idx = tid; idx < (n - c0) * kNb; idx += kLThreads)
      pan[idx] = (idx % kNb) < w ? x_load2(buf + idx) : double2{0.0, 0.0};
    __syncthreads();
    // Diagonal block (rows 0..w-1 of pan), by wave 0: L_kk in place.
    if (wave == 0) {
      for (int c = 0; c < w; ++c) {
        const double dcc = pan[c * kNb + c].x;
        const bool ok = dcc > 0.0 && dcc <= 1.7e308;
        if (!ok && lane == 0 && *fail == 0) *fail = c0 + c + 1;
        const double piv = ok ? sqrt(dcc) : __builtin_nan("");
        const double ip = 1.0 / piv;
        __builtin_amdgcn_wave_barrier();
        if (lane == 0) pan[c * kNb + c] = double2{piv, 0.0};
        for (int i = c + 1 + lane; i < w; i += 64) {
          const double2 t = pan[i * kNb + c];
          pan[i * kNb + c] = double2{t.x * ip, t.y * ip};
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // pairs (i, j), c < j <= i < w
        const int m = w - c - 1;
        for (int q = lane; q < m * m; q += 64) {
          const int i = c + 1 + q / m, j = c + 1 + q % m;
          if (j > i) continue;
          const double2 a = pan[i * kNb + c], bb = pan[j * kNb + c];
          double2 t = pan[i * kNb + j];
          t.x -= a.x * bb.x + a.y * bb.y;
          t.y -= a.y * bb.x - a.x * bb.y;
          pan[i * kNb + j] = t;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
      for (int q = lane; q < kNb * kNb; q += 64) dg[q] = pan[q];
    }
    __syncthreads();
    // Panel rows below the block: L_j = P_j L_kk^-H (forward substitution over c).
    for (int jj = w + tid; jj < n - c0; jj += kLThreads) {
      double2 l[kNb];
#pragma unroll
      for (int c = 0; c < kNb; ++c) l[c] = pan[jj * kNb + c];
#pragma unroll
      for (int c = 0; c < kNb; ++c) {
        double2 t = l[c];
#pragma unroll
        for (int m = 0; m < c; ++m) {
          const double2 d = dg[c * kNb + m];  // L_kk[c][m]; x -= l_m conj(d)
          t.x -= l[m].x * d.x + l[m].y * d.y;
          t.y -= l[m].y * d.x - l[m].x * d.y;
        }
        const double ip = c < w ? 1.0 / dg[c * kNb + c].x : 0.0;
        l[c] = double2{t.x * ip, t.y * ip};
      }
#pragma unroll
      for (int c = 0; c < kNb; ++c) pan[jj * kNb + c] = l[c];
    }
    __syncthreads();
    // Owners: final panel entries and the trailing update of their rows.
    for (int r = wave; r < nr; r += kLWaves) {
      const int i = g + r * G;
      if (i >= n) break;
      if (i < c0) continue;
      gd2* Ri = Lr + (int64_t)r * n;
      for (int c = lane; c < w; c += 64)
        if (c0 + c <= i) st2(Ri + c0 + c, pan[(i - c0) * kNb + c]);
      if (i < c0 + w) continue;
      double2 li[kNb];
#pragma unroll
      for (int c = 0; c < kNb; ++c) li[c] = pan[(i - c0) * kNb + c];
      for (int j = c0 + w + lane; j <= i; j += 64) {
        double2 t = ld2(Ri + j);
#pragma unroll
        for (int c = 0; c < kNb; ++c) {
          const double2 lj = pan[(j - c0) * kNb + c];  // t -= li_c conj(lj_c)
          t.x -= li[c].x * lj.x + li[c].y * lj.y;
          t.y -= li[c].y * lj.x - li[c].x * lj.y;
        }
        st2(Ri + j, t);
      }
    }
    __syncthreads();
  }

  // Phase 2: X = L^-1.
  for (int b = 0; b < nbk; ++b) {
    const int c0 = b * kNb, w = min(kNb, n - c0), mw = c0 + w;  // columns 0..mw-1 of R_b
    gd2* buf = xb + (int64_t)(step & 1) * xstride;
    for (int r = wave; r < nr; r += kLWaves) {
      const int i = g + r * G;
      if (i >= n) break;
      if (i < c0 || i >= c0 + w) continue;
      const int c = i - c0;
      for (int m = lane; m < mw; m += 64) x_store2(buf + (int64_t)c * n + m, ld2(Rr + (int64_t)r * n + m));
      for (int m = lane; m < w; m += 64)
        x_store2(buf + (int64_t)kNb * n + c * kNb + m, ld2(Lr + (int64_t)r * n + c0 + m));
    }
    ++step;
    l_publish(flags, step);
    l_wait(flags, G, step);
    for (int idx = tid; idx < w * mw; idx += kLThreads) {
      const int c = idx / mw, m = idx - c * mw;
      pan[c * n + m] = x_load2(buf + (int64_t)c * n + m);
    }
    for (int idx = tid; idx < kNb * kNb; idx += kLThreads) {
      const int c = idx / kNb, m = idx - c * kNb;
      dg[idx] = (c < w && m < w) ? x_load2(buf + (int64_t)kNb * n + idx) : double2{0.0, 0.0};
    }
    __syncthreads();
    // X_b = L_bb^-1 R_b, one column m per thread.
    for (int m = tid; m < mw; m += kLThreads) {
      double2 xv[kNb];
#pragma unroll
      for (int c = 0; c < kNb; ++c) {
        double2 t = c < w ? pan[c * n + m] : double2{0.0, 0.0};
#pragma unroll
        for (int q = 0; q < c; ++q) {
          const double2 d = dg[c * kNb + q];  // t -= L_bb[c][q] x_q
          t.x -= d.x * xv[q].x - d.y * xv[q].y;
          t.y -= d.x * xv[q].y + d.y * xv[q].x;
        }
        const double ip = c < w ? 1.0 / dg[c * kNb + c].x : 0.0;
        xv[c] = double2{t.x * ip, t.y * ip};
        if (c < w) pan[c * n + m] = xv[c];
      }
    }
    __syncthreads();
    if (b % G == g) {
      for (int c = 0; c < w; ++c)
        for (int m = tid; m < n; m += kLThreads)
          st2(Xo + (int64_t)(c0 + c) * n + m, m <= c0 + c ? pan[c * n + m] : double2{0.0, 0.0});
    }
    // Owners: R_i -= L_i,b X_b for their rows after the block.
    for (int r = wave; r < nr; r += kLWaves) {
      const int i = g + r * G;
      if (i >= n) break;
      if (i < c0 + w) continue;
      gd2* Ri = Rr + (int64_t)r * n;
      double2 li[kNb];
#pragma unroll
      for (int c = 0; c < kNb; ++c) li[c] = c < w ? ld2(Lr + (int64_t)r * n + c0 + c) : double2{0.0, 0.0};
      for (int m = lane; m < mw; m += 64) {
        double2 t = ld2(Ri + m);
#pragma unroll
        for (int c = 0; c < kNb; ++c) {
          const double2 xv = c < w ? pan[c * n + m] : double2{0.0, 0.0};  // t -= li_c x_c
          t.x -= li[c].x * xv.x - li[c].y * xv.y;
          t.y -= li[c].x * xv.y + li[c].y * xv.x;
        }
        st2(Ri + m, t);
      }
    }
    __syncthreads();
  }
  if (g == 0 && tid == 0 && info) *info = *fail;
}

size_t g_lds_potrf[5] = {0, 0, 0, 0, 0};

template <int kNb>
int launch_lpotrf(int slot, const double2* O, int n, int64_t ldo, double2* X, unsigned char* w,
                  const CLayout& lo, int G, int nr, int* info, hipStream_t st) {
  const size_t sh = ((size_t)n * kNb + kNb * kNb) * 16 + 16;
  QMX_HIP(set_lds_limit((const void*)k_lpotrf<kNb>, sh, &g_lds_potrf[slot]));
  int dev = 0, n_cu = 0, per_cu = 0;
  QMX_HIP(hipGetDevice(&dev));
  QMX_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
  QMX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_lpotrf<kNb>, kLThreads, sh));
  QMX_CHECK_ARG((int64_t)per_cu * n_cu >= G, "qmx_potrf_inv: grid cannot be co-resident");
  CLayout lo_arg = lo;
  void* args[] = {(void*)&O, (void*)&n, (void*)&ldo, (void*)&X, (void*)&w, (void*)&lo_arg, (void*)&nr,
                  (void*)&info};
  QMX_HIP(hipLaunchCooperativeKernel((const void*)k_lpotrf<kNb>, dim3((unsigned)G), dim3(kLThreads),
                                     args, (unsigned)sh, st));
  return QMX_OK;
}

__global__ void __launch_bounds__(64) k_czero(unsigned char* work, int64_t off) {
  QMX_GLOBAL unsigned* flags = (QMX_GLOBAL unsigned*)(work + off);
  for (int i = threadIdx.x; i < kLMaxG + 2; i += 64) flags[i] = 0u;
}

int c_nb(int64_t n) { return n <= 768 ? 12 : n <= 1152 ? 8 : n <= 1536 ? 6 : 4; }

CLayout c_layout(int64_t n, int64_t* total) {
  int64_t off = 0;
  auto take = [&](int64_t bytes) {
    const int64_t o = off;
    off += (bytes + 255) / 256 * 256;
    return o;
  };
  const int64_t nb = c_nb(n);
  CLayout L;
  L.Lrows = take((n + kLMaxG) * n * 16);
  L.Rrows = take((n + kLMaxG) * n * 16);
  L.xbuf = take(2 * (n * nb + nb * nb) * 16);
  L.flags = take((kLMaxG + 2) * 4);
  *total = off;
  return L;
}

}  // namespace

extern "C" int64_t qmx_potrf_inv_work_bytes(int64_t n) {
  if (n < 1 || n > kLMaxN) return -1;
  int64_t total = 0;
  c_layout(n, &total);
  return total;
}

extern "C" int qmx_potrf_inv(const void* O, int64_t n, int64_t ldo, void* X, int* info, void* work,
                             int64_t work_bytes, int n_groups, void* stream) {
  QMX_CHECK_ARG(O && X && work, "qmx_potrf_inv: null argument");
  QMX_CHECK_ARG(n >= 1 && n <= kLMaxN && ldo >= n, "qmx_potrf_inv: need 1 <= n <= 2048, ldo >= n");
  QMX_CHECK_ARG(work_bytes >= qmx_potrf_inv_work_bytes(n), "qmx_potrf_inv: work area too small");
  QMX_CHECK_ARG((reinterpret_cast<uintptr_t>(work) & 255) == 0,
                "qmx_potrf_inv: work area must be 256-byte aligned");
  hipStream_t st = static_cast<hipStream_t>(stream);
  int dev = 0, n_cu = 0;
  QMX_HIP(hipGetDevice(&dev));
  QMX_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
  int64_t G = n_groups > 0 ? n_groups : std::max<int64_t>(1, (n + 3) / 4);
  G = std::max<int64_t>(1, std::min<int64_t>({G, (int64_t)kLMaxG, (int64_t)n_cu, n}));
  const int64_t nr = (n + G - 1) / G;
  G = (n + nr - 1) / nr;
  int64_t total = 0;
  const CLayout lo = c_layout(n, &total);
  unsigned char* w = static_cast<unsigned char*>(work);
  hipLaunchKernelGGL(k_czero, dim3(1), dim3(64), 0, st, w, lo.flags);
  QMX_LAUNCHED();
  const double2* Op = static_cast<const double2*>(O);
  double2* Xp = static_cast<double2*>(X);
  const int ni = (int)n, Gi = (int)G, nri = (int)nr;
  switch (c_nb(n)) {
    case 12: return launch_lpotrf<12>(0, Op, ni, ldo, Xp, w, lo, Gi, nri, info, st);
    case 8: return launch_lpotrf<8>(1, Op, ni, ldo, Xp, w, lo, Gi, nri, info, st);
    case 6: return launch_lpotrf<6>(2, Op, ni, ldo, Xp, w, lo, Gi, nri, info, st);
    default: return launch_lpotrf<4>(3, Op, ni, ldo, Xp, w, lo, Gi, nri, info, st);
  }
}
