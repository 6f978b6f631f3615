// Batched Hermitian eigensolvers for small FP64 complex matrices: cyclic
// Jacobi (n <= 96) and Householder + implicit-shift QL (n <= 256), one
// workgroup per matrix, for the subspace eigensolves of Davidson and CheFSI.
#include "qmx_common.h"

using namespace qmx;

namespace {

// Strict total order for sorting eigenvalues: ascending, NaN after every
// number, equal values (and NaNs among themselves) by index. The ranks
// #{j : sorts_before(x_j, j, x_i, i)} are then a permutation of 0..n-1 for
// any input, so index arrays built from them are always fully written.
__device__ __forceinline__ bool sorts_before(double a, int ia, double b, int ib) {
  const bool na = a != a, nb = b != b;
  if (na || nb) return (na && nb) ? ia < ib : nb;
  return a < b || (a == b && ia < ib);
}



__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

}  // namespace

// ---------------------------------------------------------------------------
// Batched cyclic Jacobi for small Hermitian matrices (n <= 96): one workgroup
// per matrix, the matrix in LDS (row stride n + 1), the eigenvectors in LDS
// for n <= 70 and in a workspace above. Each sweep runs the n - 1 rounds of
// the circle ordering (pairs of a round are disjoint); a round computes all
// pair rotations J = [[cs, sn], [-sn e^-i phi, cs e^-i phi]] (c = |c| e^i phi,
// t = sign(tau)/(|tau| + sqrt(1 + tau^2)), tau = (b - a)/(2|c|)), then applies
// A <- A J and V <- V J (column phase) and A <- J^H A (row phase). Pivots with
// |c| <= 1e-17 (|a| + |b|) are zeroed; the sweeps stop when a sweep performs
// no rotation. Eigenvalues come out ascending with their vectors, from the
// lower triangle of the input as with LAPACK UPLO = L.
namespace {

constexpr int kJacThreads = 256;
constexpr int kJacMaxN = 96;
constexpr int kJacMaxNvLds = 70;

__global__ void __launch_bounds__(kJacThreads)
    k_heevj(const double2* __restrict__ A, int n, int64_t lda, int64_t stride_a,
            double* __restrict__ E, double2* __restrict__ Vout, int64_t ldv, int64_t stride_v,
            double2* __restrict__ Vws, int max_sweeps) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int ld = n + 1;
  const int m = n + (n & 1);
  const int np = m / 2;
  const bool v_in_lds = n <= kJacMaxNvLds;
  double2* a = reinterpret_cast<double2*>(smem_raw);
  double2* v = v_in_lds ? a + n * ld : Vws + (int64_t)blockIdx.x * n * ld;
  double2* rot_ph = (v_in_lds ? v + n * ld : a + n * ld);
  double* rot_cs = reinterpret_cast<double*>(rot_ph + np);
  double* rot_sn = rot_cs + np;
  int* rot_p = reinterpret_cast<int*>(rot_sn + np);
  int* rot_q = rot_p + np;
  int* n_rot = rot_q + np;
  const double2* Ab = A + (int64_t)blockIdx.x * stride_a;
  const int tid = threadIdx.x;

  for (int idx = tid; idx < n * n; idx += kJacThreads) {
    const int i = idx / n, j = idx - i * n;
    double2 x;
    if (j < i) {
      x = Ab[(int64_t)i * lda + j];
    } else if (j > i) {
      const double2 t = Ab[(int64_t)j * lda + i];
      x = double2{t.x, -t.y};
    } else {
      x = double2{Ab[(int64_t)i * lda + i].x, 0.0};
    }
    a[i * ld + j] = x;
    v[i * ld + j] = double2{i == j ? 1.0 : 0.0, 0.0};
  }
  __syncthreads();

  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    if (tid == 0) *n_rot = 0;
    __syncthreads();
    int my_rot = 0;
    for (int r = 0; r < m - 1; ++r) {
      if (tid < np) {
        int p, q;
        if (tid == 0) {
          p = r % (m - 1);
          q = m - 1;
        } else {
          p = (r + tid) % (m - 1);
          q = (r - tid + m - 1) % (m - 1);
        }
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        double cs = 1.0, sn = 0.0;
        double2 ph = double2{1.0, 0.0};
        if (q < n) {
          const double av = a[p * ld + p].x, bv = a[q * ld + q].x;
          const double2 c = a[p * ld + q];
          const double ac = sqrt(c.x * c.x + c.y * c.y);
          if (ac <= 1e-17 * (fabs(av) + fabs(bv)) || ac < 1e-300) {
            a[p * ld + q] = double2{0.0, 0.0};
            a[q * ld + p] = double2{0.0, 0.0};
            q = -1;  // no rotation
          } else {
            ph = double2{c.x / ac, c.y / ac};
            const double tau = (bv - av) / (2.0 * ac);
            const double t = fabs(tau) > 1e150 ? 0.5 / tau
                                               : (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            cs = 1.0 / sqrt(1.0 + t * t);
            sn = t * cs;
            ++my_rot;
          }
        } else {
          q = -1;
        }
        rot_p[tid] = p;
        rot_q[tid] = q;
        rot_cs[tid] = cs;
        rot_sn[tid] = sn;
        rot_ph[tid] = ph;
      }
      __syncthreads();
      // column phase: A <- A J, V <- V J
      for (int idx = tid; idx < np * n; idx += kJacThreads) {
        const int pi = idx / n, k = idx - pi * n;
        const int q = rot_q[pi];
        if (q < 0) continue;
        const int p = rot_p[pi];
        const double cs = rot_cs[pi], sn = rot_sn[pi];
        const double2 ph = rot_ph[pi];
        // sn * conj(ph) and cs * conj(ph)
        const double2 snc = double2{sn * ph.x, -sn * ph.y}, csc = double2{cs * ph.x, -cs * ph.y};
        double2 cp = a[k * ld + p], cq = a[k * ld + q];
        a[k * ld + p] = double2{cp.x * cs - (cq.x * snc.x - cq.y * snc.y),
                                cp.y * cs - (cq.x * snc.y + cq.y * snc.x)};
        a[k * ld + q] = double2{cp.x * sn + (cq.x * csc.x - cq.y * csc.y),
                                cp.y * sn + (cq.x * csc.y + cq.y * csc.x)};
        cp = v[k * ld + p];
        cq = v[k * ld + q];
        v[k * ld + p] = double2{cp.x * cs - (cq.x * snc.x - cq.y * snc.y),
                                cp.y * cs - (cq.x * snc.y + cq.y * snc.x)};
        v[k * ld + q] = double2{cp.x * sn + (cq.x * csc.x - cq.y * csc.y),
                                cp.y * sn + (cq.x * csc.y + cq.y * csc.x)};
      }
      __syncthreads();
      // row phase: A <- J^H A
      for (int idx = tid; idx < np * n; idx += kJacThreads) {
        const int pi = idx / n, k = idx - pi * n;
        const int q = rot_q[pi];
        if (q < 0) continue;
        const int p = rot_p[pi];
        const double cs = rot_cs[pi], sn = rot_sn[pi];
        const double2 ph = rot_ph[pi];
        const double2 snp = double2{sn * ph.x, sn * ph.y}, csp = double2{cs * ph.x, cs * ph.y};
        const double2 rp = a[p * ld + k], rq = a[q * ld + k];
        a[p * ld + k] = double2{cs * rp.x - (snp.x * rq.x - snp.y * rq.y),
                                cs * rp.y - (snp.x * rq.y + snp.y * rq.x)};
        a[q * ld + k] = double2{sn * rp.x + (csp.x * rq.x - csp.y * rq.y),
                                sn * rp.y + (csp.x * rq.y + csp.y * rq.x)};
      }
      __syncthreads();
    }
    if (my_rot) atomicAdd(n_rot, my_rot);
    __syncthreads();
    const int total = *n_rot;
    __syncthreads();
    if (total == 0) break;
  }

  // Ascending order: rank of each eigenvalue (ties by index).
  double* Eb = E + (int64_t)blockIdx.x * n;
  double2* Vb = Vout + (int64_t)blockIdx.x * stride_v;
  for (int i = tid; i < n; i += kJacThreads) {
    const double li = a[i * ld + i].x;
    int rank = 0;
    for (int j = 0; j < n; ++j) rank += sorts_before(a[j * ld + j].x, j, li, i);
    Eb[rank] = li;
    for (int k = 0; k < n; ++k) Vb[(int64_t)k * ldv + rank] = v[k * ld + i];
  }
}

}  // namespace

extern "C" int64_t qmx_heevj_workspace_bytes(int64_t n, int64_t batch) {
  if (n < 1 || batch < 0) return -1;
  return n <= kJacMaxNvLds ? 0 : batch * n * (n + 1) * (int64_t)sizeof(double2);
}

extern "C" int qmx_heevj_batched(const void* A, int64_t n, int64_t lda, int64_t stride_a,
                                 int64_t batch, double* E, void* V, int64_t ldv, int64_t stride_v,
                                 void* workspace, int64_t workspace_bytes, int max_sweeps,
                                 void* stream) {
  QMX_CHECK_ARG(A && E && V, "qmx_heevj_batched: null argument");
  QMX_CHECK_ARG(n >= 1 && n <= kJacMaxN, "qmx_heevj_batched: n must be in 1..96");
  QMX_CHECK_ARG(lda >= n && ldv >= n && batch >= 0 && batch < (1ll << 31),
                "qmx_heevj_batched: bad strides or batch");
  QMX_CHECK_ARG(n <= kJacMaxNvLds || (workspace && workspace_bytes >= qmx_heevj_workspace_bytes(n, batch)),
                "qmx_heevj_batched: workspace required for n > 70");
  if (batch == 0) return QMX_OK;
  const int np = (int)(n + (n & 1)) / 2;
  const size_t mats = (size_t)(n <= kJacMaxNvLds ? 2 : 1) * n * (n + 1) * sizeof(double2);
  const size_t sh = mats + (size_t)np * (sizeof(double2) + 2 * sizeof(double) + 2 * sizeof(int)) + 16;
  QMX_HIP(hipFuncSetAttribute((const void*)k_heevj, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_heevj, dim3((unsigned)batch), dim3(kJacThreads), sh, st,
                     static_cast<const double2*>(A), (int)n, lda, stride_a, E,
                     static_cast<double2*>(V), ldv, stride_v, static_cast<double2*>(workspace),
                     max_sweeps > 0 ? max_sweeps : 30);
  QMX_LAUNCHED();
  return QMX_OK;
}

// ---------------------------------------------------------------------------
// Batched Hermitian eigensolver for n <= 256, one workgroup per matrix:
//   1. Householder reduction to real tridiagonal form T = Q^H A Q (the matrix
//      lives in a workspace and stays in the cache; reflector v_k is kept in
//      column k below the subdiagonal, tau_k in LDS);
//   2. implicit-shift QL on T (tql2-style): thread 0 runs the bulge chase of
//      one QL iteration and logs its plane rotations in LDS, then every thread
//      applies the logged rotations to its own row of Z, one barrier pair per
//      QL iteration. A chase rotates the consecutive planes (m-1, m), ...,
//      (l, l+1), so the shared element of two successive rotations stays in
//      a register. Z is in LDS (row stride n + 1) for n <= kSmallLdsN and in
//      the workspace (row stride n, one row per thread, cached) above;
//   3. ascending order by rank, then V = Q Z with one thread per column,
//      applying H_{n-2} ... H_0 with the reflectors staged in LDS.
// Output as qmx_heevj_batched; info[b] != 0 if QL did not converge.
namespace {

constexpr int kSmallThreads = 256;
constexpr int kSmallMaxN = 256;
constexpr int kSmallLdsN = 128;

// Workspace bytes of one matrix: the n x n complex working copy, plus the
// n x n real QL accumulator when it does not fit in LDS; 256-byte aligned.
__host__ __device__ inline int64_t small_ws_stride(int64_t n) {
  const int64_t bytes = n * n * 16 + (n > kSmallLdsN ? n * n * 8 : 0);
  return (bytes + 255) / 256 * 256;
}

template <int NW>
__device__ __forceinline__ double block_sum_small(double x, double* red) {
  x = wave_sum_d(x);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wv] = x;
  __syncthreads();
  double s = 0;
#pragma unroll
  for (int i = 0; i < NW; ++i) s += red[i];
  return s;
}

// One matrix per workgroup: A (row stride lda, lower triangle read) ->
// E ascending, V (row stride ldv), work = small_ws_stride(n) bytes.
template <bool ZLDS>
__device__ void heev_small_one(const double2* __restrict__ Ab, int n, int64_t lda,
                               double* __restrict__ Eb, double2* __restrict__ Vb, int64_t ldv,
                               double2* __restrict__ W, int max_iter, int* __restrict__ infob) {
  constexpr int NW = kSmallThreads / 64;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int ldz = ZLDS ? n + 1 : n;
  double2* v = reinterpret_cast<double2*>(smem_raw);
  double2* w = v + n;
  double2* tau = w + n;
  double* Zl = reinterpret_cast<double*>(tau + n);
  double* d = Zl + (ZLDS ? n * ldz : 0);
  double* e = d + n;
  double* lc = e + n;   // [2][n] rotation cosines, double-buffered log
  double* ls = lc + 2 * n;  // [2][n] rotation sines
  double* red = ls + 2 * n;
  int* perm = reinterpret_cast<int*>(red + NW);
  int* ctl = perm + n;  // [0..1] log lengths, [2] failure, [3..4] last-log flags, [5..6] top plane
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // Global Z: n * n doubles right after this matrix's n * n complex W.
  double* Z = ZLDS ? Zl : reinterpret_cast<double*>(W + (int64_t)n * n);

  for (int idx = tid; idx < n * n; idx += kSmallThreads) {
    const int i = idx / n, j = idx - i * n;
    double2 x;
    if (j < i) {
      x = Ab[(int64_t)i * lda + j];
    } else if (j > i) {
      const double2 t = Ab[(int64_t)j * lda + i];
      x = double2{t.x, -t.y};
    } else {
      x = double2{Ab[(int64_t)i * lda + i].x, 0.0};
    }
    W[(int64_t)i * n + j] = x;
  }
  if (tid == 0) ctl[2] = 0;
  __syncthreads();

  // 1. Householder tridiagonalisation.
  for (int k = 0; k + 1 < n; ++k) {
    double s = 0;
    for (int i = k + 2 + tid; i < n; i += kSmallThreads) {
      const double2 x = W[(int64_t)i * n + k];
      s += x.x * x.x + x.y * x.y;
    }
    const double sigma = block_sum_small<NW>(s, red);
    const double2 alpha = W[(int64_t)(k + 1) * n + k];
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
    for (int i = k + 2 + tid; i < n; i += kSmallThreads) {
      const double2 x = W[(int64_t)i * n + k];
      const double2 vv = double2{x.x * sr - x.y * si, x.x * si + x.y * sr};
      v[i] = vv;
      W[(int64_t)i * n + k] = vv;
    }
    if (tid == 0) {
      v[k + 1] = double2{1.0, 0.0};
      tau[k] = tk;
      e[k] = beta;
      d[k] = W[(int64_t)k * n + k].x;
    }
    __syncthreads();
    if (tk.x == 0.0 && tk.y == 0.0) continue;
    // y = W[k+1:, k+1:] v, one thread per entry: W is Hermitian, so
    // y_i = sum_j conj(W[j][i]) v_j reads column i (coalesced over i).
    for (int i = k + 1 + tid; i < n; i += kSmallThreads) {
      double yr = 0, yi = 0;
#pragma unroll 16
      for (int j = k + 1; j < n; ++j) {
        const double2 a = W[(int64_t)j * n + i], bb = v[j];
        yr += a.x * bb.x + a.y * bb.y;
        yi += a.x * bb.y - a.y * bb.x;
      }
      w[i] = double2{yr, yi};
    }
    __syncthreads();
    double part = 0;
    for (int i = k + 1 + tid; i < n; i += kSmallThreads) {
      const double2 vi = v[i], yi = w[i];
      part += vi.x * yi.x + vi.y * yi.y;
    }
    const double vy = block_sum_small<NW>(part, red);
    const double t2 = tk.x * tk.x + tk.y * tk.y;
    for (int i = k + 1 + tid; i < n; i += kSmallThreads) {
      const double2 y = w[i], vi = v[i];
      w[i] = double2{tk.x * y.x - tk.y * y.y - 0.5 * t2 * vy * vi.x,
                     tk.x * y.y + tk.y * y.x - 0.5 * t2 * vy * vi.y};
    }
    __syncthreads();
    const int m = n - k - 1;
    for (int idx = tid; idx < m * m; idx += kSmallThreads) {
      const int i = k + 1 + idx / m, j = k + 1 + idx % m;
      const double2 vi = v[i], wi = w[i], vj = v[j], wj = w[j];
      double2 a = W[(int64_t)i * n + j];
      a.x -= vi.x * wj.x + vi.y * wj.y + wi.x * vj.x + wi.y * vj.y;
      a.y -= vi.y * wj.x - vi.x * wj.y + wi.y * vj.x - wi.x * vj.y;
      W[(int64_t)i * n + j] = a;
    }
    __syncthreads();
  }
  if (tid == 0) {
    d[n - 1] = W[(int64_t)(n - 1) * n + (n - 1)].x;
    e[n - 1] = 0.0;
    if (n == 1) tau[0] = double2{0.0, 0.0};
  }
  for (int idx = tid; idx < n * ldz; idx += kSmallThreads) {
    const int i = idx / ldz, j = idx - i * ldz;
    Z[idx] = (i == j) ? 1.0 : 0.0;
  }
  if (!ZLDS) __threadfence_block();  // global Z written by other threads of the group
  __syncthreads();

  // 2. Implicit-shift QL (tql2 order: eigenvalue l = 0, 1, ... deflates
  // when e[l] is negligible). Lane 0 of wave 0 runs the bulge chase of one
  // QL iteration into a rotation log while waves 1-3 apply the previous log
  // to their rows of Z (double-buffered logs, one barrier per iteration).
  // A chase rotates the consecutive planes (ihi, ihi + 1), (ihi - 1, ihi),
  // ...; each row applies them in blocks of 8 with the reads hoisted, the
  // element shared by successive rotations carried in a register.
  {
    const double eps = 2.220446049250313e-16;
    int l = 0, iter = 0;  // chase state, lane 0 of wave 0 only
    // Produce the next log into buffer `bf`: ctl[bf] rotations, ctl[3 + bf]
    // set when no rotations remain.
    auto produce = [&](int bf) {
      double* cl = lc + bf * n;
      double* sl = ls + bf * n;
      for (;;) {
        if (l >= n) {
          ctl[bf] = 0;
          ctl[3 + bf] = 1;
          return;
        }
        int m = l;
        while (m < n - 1) {
          const double dd = fabs(d[m]) + fabs(d[m + 1]);
          if (fabs(e[m]) <= eps * dd) break;
          ++m;
        }
        if (m == l) {
          ++l;
          iter = 0;
          continue;
        }
        if (++iter > max_iter) {
          ctl[2] = 1;
          ++l;
          iter = 0;
          continue;
        }
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = sqrt(g * g + 1.0);
        g = d[m] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
        double s = 1.0, c = 1.0, p = 0.0;
        int nlog = 0;
        bool underflow = false;
        // Step i reads d[i], e[i] and d[i + 1] before any step writes them
        // (step i writes d[i + 1] and e[i + 1] only, after step i + 1 read
        // d[i + 1] as its d[i]), so they stream through registers.
        double di = d[m - 1], ei = e[m - 1], dip1 = d[m];
        for (int i = m - 1; i >= l; --i) {
          const double dn = i > l ? d[i - 1] : 0.0, en = i > l ? e[i - 1] : 0.0;
          const double f = s * ei, bb = c * ei;
          const double x = f * f + g * g;
          if (x == 0.0) {
            e[i + 1] = 0.0;
            d[i + 1] = dip1 - p;
            e[m] = 0.0;
            underflow = true;
            break;
          }
          // 1/sqrt(x): hardware estimate and two Newton steps (full FP64)
          double ri = __builtin_amdgcn_rsq(x);
          ri = ri * (1.5 - 0.5 * x * ri * ri);
          ri = ri * (1.5 - 0.5 * x * ri * ri);
          r = x * ri;
          e[i + 1] = r;
          s = f * ri;
          c = g * ri;
          g = dip1 - p;
          r = (di - g) * s + 2.0 * c * bb;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - bb;
          cl[nlog] = c;
          sl[nlog] = s;
          ++nlog;
          dip1 = di;
          di = dn;
          ei = en;
        }
        if (!underflow) {
          d[l] -= p;
          e[l] = g;
          e[m] = 0.0;
        }
        ctl[bf] = nlog;
        ctl[3 + bf] = 0;
        ctl[5 + bf] = m - 1;
        return;
      }
    };
    if (tid == 0) produce(0);
    __syncthreads();
    for (int step = 0;; ++step) {
      const int cur = step & 1;
      const int nlog = ctl[cur];
      if (ctl[3 + cur]) break;  // no log left (every row is up to date)
      if (wave == 0) {
        if (lane == 0) produce(cur ^ 1);
      } else if (nlog > 0) {
        const double* cl = lc + cur * n;
        const double* sl = ls + cur * n;
        const int ihi = ctl[5 + cur];
        for (int row = tid - 64; row < n; row += kSmallThreads - 64) {
          double* zr = Z + (int64_t)row * ldz;
          double carry = zr[ihi + 1];
          for (int t0 = 0; t0 < nlog; t0 += 8) {
            const int cnt = min(8, nlog - t0);
            double zz[8];
#pragma unroll
            for (int q = 0; q < 8; ++q)
              if (q < cnt) zz[q] = zr[ihi - t0 - q];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
              if (q < cnt) {
                const double c = cl[t0 + q], s = sl[t0 + q];
                zr[ihi - t0 - q + 1] = s * zz[q] + c * carry;
                carry = c * zz[q] - s * carry;
              }
            }
          }
          zr[ihi - nlog + 1] = carry;
        }
      }
      __syncthreads();
    }
  }

  // 3. Ascending order and V = Q Z.
  for (int i = tid; i < n; i += kSmallThreads) {
    const double li_ = d[i];
    int rank = 0;
    for (int j = 0; j < n; ++j) rank += sorts_before(d[j], j, li_, i);
    perm[rank] = i;
    Eb[rank] = li_;
  }
  if (tid == 0 && infob) *infob = ctl[2];
  __syncthreads();
  for (int j = tid; j < n; j += kSmallThreads) {
    const int pj = perm[j];
    for (int i = 0; i < n; ++i) Vb[(int64_t)i * ldv + j] = double2{Z[(int64_t)i * ldz + pj], 0.0};
  }
  for (int k = n - 2; k >= 0; --k) {
    __syncthreads();
    for (int i = k + 2 + tid; i < n; i += kSmallThreads) v[i] = W[(int64_t)i * n + k];
    if (tid == 0) v[k + 1] = double2{1.0, 0.0};
    __syncthreads();
    const double2 tk = tau[k];
    if (tk.x == 0.0 && tk.y == 0.0) continue;
    for (int j = tid; j < n; j += kSmallThreads) {
      double sr = 0, si = 0;  // s = v^H V[:, j]
      for (int i = k + 1; i < n; ++i) {
        const double2 vi = v[i], x = Vb[(int64_t)i * ldv + j];
        sr += vi.x * x.x + vi.y * x.y;
        si += vi.x * x.y - vi.y * x.x;
      }
      const double ts_r = tk.x * sr - tk.y * si, ts_i = tk.x * si + tk.y * sr;  // tau * s
      for (int i = k + 1; i < n; ++i) {
        const double2 vi = v[i];
        double2 x = Vb[(int64_t)i * ldv + j];
        x.x -= vi.x * ts_r - vi.y * ts_i;
        x.y -= vi.x * ts_i + vi.y * ts_r;
        Vb[(int64_t)i * ldv + j] = x;
      }
    }
  }
}

template <bool ZLDS>
__global__ void __launch_bounds__(kSmallThreads)
    k_heev_small(const double2* __restrict__ A, int n, int64_t lda, int64_t stride_a,
                 double* __restrict__ E, double2* __restrict__ Vout, int64_t ldv,
                 int64_t stride_v, double2* __restrict__ Wws, int max_iter, int* __restrict__ info) {
  const int64_t b = blockIdx.x;
  heev_small_one<ZLDS>(A + b * stride_a, n, lda, E + b * n, Vout + b * stride_v, ldv,
                       reinterpret_cast<double2*>(reinterpret_cast<char*>(Wws) + b * small_ws_stride(n)),
                       max_iter, info ? info + b : nullptr);
}

// Matrices of different sizes in one launch, one descriptor each.
template <bool ZLDS>
__global__ void __launch_bounds__(kSmallThreads)
    k_heev_small_grouped(const qmx_heev_desc* __restrict__ desc, int max_iter) {
  const qmx_heev_desc d = desc[blockIdx.x];
  heev_small_one<ZLDS>(static_cast<const double2*>(d.A), d.n, d.lda, d.E,
                       static_cast<double2*>(d.V), d.ldv, static_cast<double2*>(d.work), max_iter,
                       d.info);
}

static size_t small_lds_bytes(int64_t n) {
  const bool zlds = n <= kSmallLdsN;
  return (size_t)3 * n * sizeof(double2) + (zlds ? (size_t)n * (n + 1) * sizeof(double) : 0) +
         (size_t)6 * n * sizeof(double) + 4 * sizeof(double) + (size_t)n * sizeof(int) +
         8 * sizeof(int);
}

}  // namespace

extern "C" int64_t qmx_heev_small_workspace_bytes(int64_t n, int64_t batch) {
  if (n < 1 || batch < 0) return -1;
  return batch * small_ws_stride(n);
}

extern "C" int qmx_heev_small_batched(const void* A, int64_t n, int64_t lda, int64_t stride_a,
                                      int64_t batch, double* E, void* V, int64_t ldv,
                                      int64_t stride_v, int* info, void* workspace,
                                      int64_t workspace_bytes, void* stream) {
  QMX_CHECK_ARG(A && E && V && workspace, "qmx_heev_small_batched: null argument");
  QMX_CHECK_ARG(n >= 1 && n <= kSmallMaxN, "qmx_heev_small_batched: n must be in 1..256");
  QMX_CHECK_ARG(lda >= n && ldv >= n && batch >= 0 && batch < (1ll << 31),
                "qmx_heev_small_batched: bad strides or batch");
  QMX_CHECK_ARG(workspace_bytes >= qmx_heev_small_workspace_bytes(n, batch),
                "qmx_heev_small_batched: workspace too small");
  if (batch == 0) return QMX_OK;
  const bool zlds = n <= kSmallLdsN;
  const size_t sh = small_lds_bytes(n);
  const void* fn = zlds ? (const void*)k_heev_small<true> : (const void*)k_heev_small<false>;
  QMX_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (zlds)
    hipLaunchKernelGGL(k_heev_small<true>, dim3((unsigned)batch), dim3(kSmallThreads), sh, st,
                       static_cast<const double2*>(A), (int)n, lda, stride_a, E,
                       static_cast<double2*>(V), ldv, stride_v, static_cast<double2*>(workspace),
                       60, info);
  else
    hipLaunchKernelGGL(k_heev_small<false>, dim3((unsigned)batch), dim3(kSmallThreads), sh, st,
                       static_cast<const double2*>(A), (int)n, lda, stride_a, E,
                       static_cast<double2*>(V), ldv, stride_v, static_cast<double2*>(workspace),
                       60, info);
  QMX_LAUNCHED();
  return QMX_OK;
}

extern "C" int64_t qmx_heev_small_work_bytes(int64_t n) {
  if (n < 1 || n > kSmallMaxN) return -1;
  return (n * n * 24 + 255) / 256 * 256;  // complex copy + real accumulator, any launch variant
}

extern "C" int qmx_heev_small_grouped(const qmx_heev_desc* desc, int64_t count, int64_t n_max,
                                      void* stream) {
  QMX_CHECK_ARG(desc || count == 0, "qmx_heev_small_grouped: null descriptors");
  QMX_CHECK_ARG(count >= 0 && count < (1ll << 31), "qmx_heev_small_grouped: bad count");
  QMX_CHECK_ARG(n_max >= 1 && n_max <= kSmallMaxN, "qmx_heev_small_grouped: n_max must be in 1..256");
  if (count == 0) return QMX_OK;
  const bool zlds = n_max <= kSmallLdsN;
  const size_t sh = small_lds_bytes(n_max);
  const void* fn = zlds ? (const void*)k_heev_small_grouped<true> : (const void*)k_heev_small_grouped<false>;
  QMX_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (zlds)
    hipLaunchKernelGGL(k_heev_small_grouped<true>, dim3((unsigned)count), dim3(kSmallThreads), sh,
                       st, desc, 60);
  else
    hipLaunchKernelGGL(k_heev_small_grouped<false>, dim3((unsigned)count), dim3(kSmallThreads),
                       sh, st, desc, 60);
  QMX_LAUNCHED();
  return QMX_OK;
}
