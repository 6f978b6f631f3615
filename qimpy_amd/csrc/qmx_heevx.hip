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
