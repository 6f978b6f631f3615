// Band-wise elementwise and reduction kernels of the Davidson loop (gfx950).
//
// All are HBM-bound streams over (k, band, g) rows: one workgroup per row,
// accumulating in double with a wave64 shuffle reduction.
#include <type_traits>

#include "qmx_common.h"

using namespace qmx;

namespace {

constexpr int kRowThreads = 256;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Block-wide sum of two values; result valid in thread 0.
__device__ __forceinline__ void block_sum2(double& a, double& b) {
  __shared__ double sa[kRowThreads / 64], sb[kRowThreads / 64];
  a = wave_sum(a);
  b = wave_sum(b);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    sa[w] = a;
    sb[w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    a = 0;
    b = 0;
    for (int i = 0; i < kRowThreads / 64; ++i) {
      a += sa[i];
      b += sb[i];
    }
  }
}

template <typename C2>
__global__ void __launch_bounds__(kRowThreads)
    k_band_reduce(const C2* __restrict__ C, int64_t ldb, int64_t ldk, int nb, int n_tot,
                  const double* __restrict__ ke, int64_t ke_ldk, double* __restrict__ norm2,
                  double* __restrict__ kin) {
  const int b = blockIdx.x, k = blockIdx.y;
  const C2* row = C + k * ldk + (int64_t)b * ldb;
  const double* kk = ke ? ke + k * ke_ldk : nullptr;
  double s = 0, e = 0;
  for (int g = threadIdx.x; g < n_tot; g += blockDim.x) {
    const C2 c = row[g];
    const double m = (double)c.x * c.x + (double)c.y * c.y;
    s += m;
    if (kk) e += m * kk[g];
  }
  block_sum2(s, e);
  if (threadIdx.x == 0) {
    norm2[(int64_t)k * nb + b] = s;
    if (kin) kin[(int64_t)k * nb + b] = e;
  }
}

// The preconditioner factor 1 / (x + exp(-x)), x = KE / KEref, in FP64, or
// in FP32 (LOW) where the residual itself is stored in float32. x is formed
// in FP64 first: 1 / KEref of a (near-)null band can exceed the float range,
// and KE = 0 times an infinite float would give NaN where FP64 gives 0; a
// float x that overflows to inf gives the factor 0, as FP64 does.
template <bool LOW>
__device__ __forceinline__ double precond_factor(double ke, double inv_ref) {
  if constexpr (LOW) {
    const float x = (float)(ke * inv_ref);
    return (double)(1.0f / (x + __expf(-x)));
  } else {
    const double x = ke * inv_ref;
    return 1.0 / (x + exp(-x));
  }
}

// NORM: the row is then scaled to unit norm (zero where its norm is below
// 1e-30) and its padding g >= nvec[k] zeroed, in the same launch (the row
// is still in cache when it is re-read).
template <typename C2, bool LOW, bool NORM>
__global__ void __launch_bounds__(kRowThreads)
    k_precond_residual(const C2* __restrict__ C, const C2* __restrict__ HC, int64_t ldb,
                       int64_t ldk, const double* __restrict__ eig,
                       const double* __restrict__ KEref, const double* __restrict__ ke,
                       int64_t ke_ldk, C2* __restrict__ out, int64_t oldb, int64_t oldk, int nb,
                       int n_tot, double* __restrict__ norm2, double* __restrict__ rnorm2,
                       const int* __restrict__ nvec) {
  using R = decltype(C2{}.x);
  const int b = blockIdx.x, k = blockIdx.y;
  const int64_t ir = (int64_t)k * nb + b;
  const C2* c = C + k * ldk + (int64_t)b * ldb;
  const C2* h = HC + k * ldk + (int64_t)b * ldb;
  C2* o = out + k * oldk + (int64_t)b * oldb;
  const double* kk = ke + k * ke_ldk;
  const double e = eig[ir];
  double ref;
  if (KEref) {
    ref = KEref[ir];
  } else {  // the band's kinetic energy sum_g |C|^2 KE first (the row stays in cache)
    double kin = 0, unused = 0;
    for (int g = threadIdx.x; g < n_tot; g += blockDim.x) {
      const C2 cc = c[g];
      kin += ((double)cc.x * cc.x + (double)cc.y * cc.y) * kk[g];
    }
    block_sum2(kin, unused);
    __shared__ double kin_all;
    if (threadIdx.x == 0) kin_all = kin;
    __syncthreads();
    ref = kin_all;
  }
  const double inv_ref = 1.0 / ref;
  double s = 0, sr = 0;
  for (int g = threadIdx.x; g < n_tot; g += blockDim.x) {
    const double inv_x = precond_factor<LOW>(kk[g], inv_ref);
    const C2 cc = c[g], hh = h[g];
    const double ex = (double)hh.x - e * (double)cc.x;
    const double ey = (double)hh.y - e * (double)cc.y;
    sr += ex * ex + ey * ey;
    const C2 r = C2{(R)(ex * inv_x), (R)(ey * inv_x)};
    o[g] = r;
    s += (double)r.x * r.x + (double)r.y * r.y;
  }
  block_sum2(s, sr);
  if (threadIdx.x == 0) {
    norm2[ir] = s;
    if (rnorm2) rnorm2[ir] = sr;
  }
  if constexpr (NORM) {
    __shared__ double scale;
    if (threadIdx.x == 0) scale = s > 1e-60 ? 1.0 / sqrt(s) : 0.0;
    __syncthreads();
    const R f = (R)scale;
    const int nk_valid = nvec[k];
    for (int g = threadIdx.x; g < n_tot; g += blockDim.x) {
      if (g >= nk_valid || f == R(0)) {  // a null (or non-finite) row becomes zero
        o[g] = C2{R(0), R(0)};
      } else {
        const C2 v = o[g];
        o[g] = C2{v.x * f, v.y * f};
      }
    }
  }
}

template <typename C2>
__global__ void __launch_bounds__(kRowThreads)
    k_scale_constrain(C2* __restrict__ C, int64_t ldb, int64_t ldk, int nb, int n_tot,
                      const double* __restrict__ scale, const int* __restrict__ nvec) {
  using R = decltype(C2{}.x);
  const int b = blockIdx.x, k = blockIdx.y;
  C2* row = C + k * ldk + (int64_t)b * ldb;
  const R s = scale ? (R)scale[(int64_t)k * nb + b] : R(1);
  const int nk_valid = nvec[k];
  for (int g = threadIdx.x; g < n_tot; g += blockDim.x) {
    if (g >= nk_valid) {
      row[g] = C2{0, 0};
    } else if (scale) {
      const C2 v = row[g];
      row[g] = C2{v.x * s, v.y * s};
    }
  }
}

// Signed int64 xorshift step with arithmetic right shift, as on torch
// int64 tensors; returns 2^-64 x + 1/2 in double.
__device__ __forceinline__ double xorshift_u(int64_t& x) {
  x ^= (int64_t)((uint64_t)x << 13);
  x ^= x >> 7;
  x ^= (int64_t)((uint64_t)x << 17);
  return 5.42101086242752217e-20 * (double)x + 0.5;
}

template <typename C2>
__global__ void k_randomize(C2* __restrict__ C, int64_t ldb, int64_t ldk, int n_tot,
                            int64_t n_max, int64_t k_offset, const int* __restrict__ nvec,
                            const double* __restrict__ ke, int64_t ke_ldk, int64_t seed,
                            int64_t b_start, int64_t b_stop) {
  using R = decltype(C2{}.x);
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (g >= n_tot) return;
  int64_t x = (1 + seed) + g + n_max * (k_offset + k);
  for (int64_t i = 0; i <= b_start; ++i) {
    xorshift_u(x);
    xorshift_u(x);
  }
  const bool pad = g >= nvec[k];
  const double e = ke[k * ke_ldk + g];
  const double t = (4.0 / 3) * e;
  const double t2 = t * t;
  const double limit = 1.0 / (1.0 + t2 * t2 * t2);
  for (int64_t b = b_start; b < b_stop; ++b) {
    const double u0 = xorshift_u(x);
    const double u1 = xorshift_u(x);
    const double mag = sqrt(-2.0 * log(u0));
    double sn, cs;
    sincos(2.0 * 3.141592653589793 * u1, &sn, &cs);
    C2 v = pad ? C2{0, 0} : C2{(R)(mag * cs * limit), (R)(mag * sn * limit)};
    C[k * ldk + b * ldb + g] = v;
  }
}

}  // namespace

namespace {

// out[k][b][g] = [accumulate ? out : 0] + alpha i G[k][g][dir] C[k][b][g]:
// the gradient operator of Basis.apply_gradient (dft/electrons/
// _basis_ops.py:15-26; get_gradient = i (iG + k) Gbasis^T, _basis.py:240-254),
// scaled and optionally accumulated. One thread per plane wave; out may be C.
template <typename C2, typename R>
__global__ void __launch_bounds__(256)
    k_apply_gradient(const C2* C, int64_t ldb, int64_t ldk, C2* out, int64_t oldb, int64_t oldk,
                     int n_tot, const R* __restrict__ G, int dir, R alpha, int accumulate) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  const int64_t b = blockIdx.y, k = blockIdx.z;
  if (g >= n_tot) return;
  const R s = alpha * G[(k * n_tot + g) * 3 + dir];
  const C2 c = C[k * ldk + b * ldb + g];
  C2 v{-s * c.y, s * c.x};  // i s c
  C2* o = out + k * oldk + b * oldb + g;
  if (accumulate) {
    const C2 prev = *o;
    v.x += prev.x;
    v.y += prev.y;
  }
  *o = v;
}

// out = a X + b Y + c Z on (k, band, g) rows, a, b, c per (k, band) in
// double (c, Z optional); one workgroup per row, out may alias an input.
template <typename C2>
__global__ void __launch_bounds__(kRowThreads)
    k_axpby(const double* __restrict__ ca, const double* __restrict__ cb,
            const double* __restrict__ cc, int64_t cf_ldk, int64_t cf_ldb, const C2* X,
            const C2* Y, const C2* Z, C2* out, int64_t ldb, int64_t ldk, int n_tot) {
  using R = decltype(C2{}.x);
  const int b = blockIdx.x, k = blockIdx.y;
  const int64_t cf = k * cf_ldk + (int64_t)b * cf_ldb;
  const int64_t row = k * ldk + (int64_t)b * ldb;
  const R a = (R)ca[cf], bb = cb ? (R)cb[cf] : R(0), c = (cc && Z) ? (R)cc[cf] : R(0);
  for (int g = threadIdx.x; g < n_tot; g += blockDim.x) {
    const C2 x = X[row + g];
    C2 v{a * x.x, a * x.y};
    if (cb) {
      const C2 y = Y[row + g];
      v.x += bb * y.x;
      v.y += bb * y.y;
    }
    if (cc && Z) {
      const C2 z = Z[row + g];
      v.x += c * z.x;
      v.y += c * z.y;
    }
    out[row + g] = v;
  }
}

// Row blocks between device buffers in one launch: block i writes
// dst[i0 * dst_s0 + i1 * dst_s1 + c] for i0 < n0, i1 < n1, c < dst_cols,
// from src at the same (i0, i1) with src strides for c < src_cols and zero
// beyond (or everywhere when src is null). mode 0 copies 8-byte words
// (16-byte ones when every pointer, stride and width allows); mode 1
// converts doubles to floats (counts and strides in reals of each side).
// Workgroups are dealt to blocks in order: block i owns first_wg onwards,
// rows_per_wg rows of its n0 x n1 rows each.
struct CopyBlock {
  const void* src;
  void* dst;
  int64_t n0, n1, src_cols, dst_cols, src_s0, src_s1, dst_s0, dst_s1, rows_per_wg, first_wg,
      mode, pad;
};
static_assert(sizeof(CopyBlock) == 14 * 8, "CopyBlock is 14 int64 fields");

template <typename S, typename D>
__device__ __forceinline__ D convert_unit(S v) {
  if constexpr (std::is_same<S, D>::value)
    return v;
  else
    return (D)v;
}

template <typename S, typename D>
__device__ __forceinline__ void copy_rows(const CopyBlock& b, int64_t r0, int64_t r1, int u) {
  const int64_t sc = b.src_cols / u, dc = b.dst_cols / u;
  for (int64_t r = r0; r < r1; ++r) {
    const int64_t i0 = r / b.n1, i1 = r - i0 * b.n1;
    const S* s = b.src ? static_cast<const S*>(b.src) + (i0 * b.src_s0 + i1 * b.src_s1) / u
                       : nullptr;
    D* o = static_cast<D*>(b.dst) + (i0 * b.dst_s0 + i1 * b.dst_s1) / u;
    for (int64_t c = threadIdx.x; c < dc; c += kRowThreads) {
      D v;
      if (s && c < sc)
        v = convert_unit<S, D>(s[c]);
      else
        v = D{};
      o[c] = v;
    }
  }
}

__global__ void __launch_bounds__(kRowThreads)
    k_copy_blocks(const CopyBlock* __restrict__ blocks, int n) {
  const int64_t wg = blockIdx.x;
  int lo = 0, hi = n - 1;  // the last block with first_wg <= wg
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (blocks[mid].first_wg <= wg)
      lo = mid;
    else
      hi = mid - 1;
  }
  const CopyBlock b = blocks[lo];
  const int64_t rows = b.n0 * b.n1;
  const int64_t r0 = (wg - b.first_wg) * b.rows_per_wg;
  const int64_t r1 = r0 + b.rows_per_wg < rows ? r0 + b.rows_per_wg : rows;
  if (b.mode == 1) {
    copy_rows<double, float>(b, r0, r1, 1);
    return;
  }
  const bool wide = ((b.src_cols | b.dst_cols | b.src_s0 | b.src_s1 | b.dst_s0 | b.dst_s1) & 1) == 0 &&
                    ((reinterpret_cast<uintptr_t>(b.src) | reinterpret_cast<uintptr_t>(b.dst)) & 15) == 0;
  if (wide)
    copy_rows<ulonglong2, ulonglong2>(b, r0, r1, 2);
  else
    copy_rows<unsigned long long, unsigned long long>(b, r0, r1, 1);
}

}  // namespace

#define QMX_DISPATCH(dtype, KERNEL_CALL_F32, KERNEL_CALL_F64)          \
  do {                                                                \
    if ((dtype) == QMX_C64) {                                         \
      KERNEL_CALL_F32;                                                \
    } else if ((dtype) == QMX_C128) {                                 \
      KERNEL_CALL_F64;                                                \
    } else {                                                          \
      qmx::set_error("dtype must be QMX_C64 or QMX_C128");            \
      return QMX_ERR_ARG;                                             \
    }                                                                 \
    QMX_LAUNCHED();                                                   \
  } while (0)

extern "C" {

int qmx_band_reduce(int dtype, const void* C, int64_t ldb, int64_t ldk, int64_t nk, int64_t nb,
                    int64_t n_tot, const double* ke, int64_t ke_ldk, double* norm2, double* kin,
                    void* stream) {
  QMX_CHECK_ARG(C && norm2, "qmx_band_reduce: null argument");
  QMX_CHECK_ARG(nk >= 0 && nk < 65536 && nb >= 0 && nb < (1 << 30) && n_tot >= 0,
                "qmx_band_reduce: bad sizes");
  QMX_CHECK_ARG(!kin || ke, "qmx_band_reduce: kin requires ke");
  if (nk == 0 || nb == 0) return QMX_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  dim3 grid((unsigned)nb, (unsigned)nk);
  QMX_DISPATCH(dtype,
               hipLaunchKernelGGL(k_band_reduce<float2>, grid, dim3(kRowThreads), 0, st,
                                  (const float2*)C, ldb, ldk, (int)nb, (int)n_tot, ke, ke_ldk,
                                  norm2, kin),
               hipLaunchKernelGGL(k_band_reduce<double2>, grid, dim3(kRowThreads), 0, st,
                                  (const double2*)C, ldb, ldk, (int)nb, (int)n_tot, ke, ke_ldk,
                                  norm2, kin));
  return QMX_OK;
}

int qmx_precond_residual(int dtype, const void* C, const void* HC, int64_t ldb, int64_t ldk,
                         const double* eig, const double* KEref, const double* ke,
                         int64_t ke_ldk, void* out, int64_t out_ldb, int64_t out_ldk,
                         int64_t nk, int64_t nb, int64_t n_tot, double* norm2, double* rnorm2,
                         void* stream) {
  QMX_CHECK_ARG(C && HC && eig && ke && out && norm2,
                "qmx_precond_residual: null argument");
  QMX_CHECK_ARG(nk >= 0 && nk < 65536 && nb >= 0 && nb < (1 << 30) && n_tot >= 0,
                "qmx_precond_residual: bad sizes");
  if (nk == 0 || nb == 0) return QMX_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  dim3 grid((unsigned)nb, (unsigned)nk);
  QMX_DISPATCH(dtype,
               hipLaunchKernelGGL((k_precond_residual<float2, true, false>), grid,
                                  dim3(kRowThreads), 0, st, (const float2*)C, (const float2*)HC,
                                  ldb, ldk, eig, KEref, ke, ke_ldk, (float2*)out, out_ldb,
                                  out_ldk, (int)nb, (int)n_tot, norm2, rnorm2, nullptr),
               hipLaunchKernelGGL((k_precond_residual<double2, false, false>), grid,
                                  dim3(kRowThreads), 0, st, (const double2*)C, (const double2*)HC,
                                  ldb, ldk, eig, KEref, ke, ke_ldk, (double2*)out, out_ldb,
                                  out_ldk, (int)nb, (int)n_tot, norm2, rnorm2, nullptr));
  return QMX_OK;
}

int qmx_precond_residual_normalize(int dtype, const void* C, const void* HC, int64_t ldb,
                                   int64_t ldk, const double* eig, const double* KEref,
                                   const double* ke, int64_t ke_ldk, void* out, int64_t out_ldb,
                                   int64_t out_ldk, int64_t nk, int64_t nb, int64_t n_tot,
                                   double* norm2, double* rnorm2, const int32_t* n,
                                   void* stream) {
  QMX_CHECK_ARG(C && HC && eig && KEref && ke && out && norm2 && n,
                "qmx_precond_residual_normalize: null argument");
  QMX_CHECK_ARG(nk >= 0 && nk < 65536 && nb >= 0 && nb < (1 << 30) && n_tot >= 0,
                "qmx_precond_residual_normalize: bad sizes");
  if (nk == 0 || nb == 0) return QMX_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  dim3 grid((unsigned)nb, (unsigned)nk);
  QMX_DISPATCH(dtype,
               hipLaunchKernelGGL((k_precond_residual<float2, true, true>), grid,
                                  dim3(kRowThreads), 0, st, (const float2*)C, (const float2*)HC,
                                  ldb, ldk, eig, KEref, ke, ke_ldk, (float2*)out, out_ldb,
                                  out_ldk, (int)nb, (int)n_tot, norm2, rnorm2, n),
               hipLaunchKernelGGL((k_precond_residual<double2, false, true>), grid,
                                  dim3(kRowThreads), 0, st, (const double2*)C, (const double2*)HC,
                                  ldb, ldk, eig, KEref, ke, ke_ldk, (double2*)out, out_ldb,
                                  out_ldk, (int)nb, (int)n_tot, norm2, rnorm2, n));
  return QMX_OK;
}

int qmx_scale_constrain(int dtype, void* C, int64_t ldb, int64_t ldk, int64_t nk, int64_t nb,
                        int64_t n_tot, const double* scale, const int32_t* n, void* stream) {
  QMX_CHECK_ARG(C && n, "qmx_scale_constrain: null argument");
  QMX_CHECK_ARG(nk >= 0 && nk < 65536 && nb >= 0 && nb < (1 << 30) && n_tot >= 0,
                "qmx_scale_constrain: bad sizes");
  if (nk == 0 || nb == 0) return QMX_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  dim3 grid((unsigned)nb, (unsigned)nk);
  QMX_DISPATCH(dtype,
               hipLaunchKernelGGL(k_scale_constrain<float2>, grid, dim3(kRowThreads), 0, st,
                                  (float2*)C, ldb, ldk, (int)nb, (int)n_tot, scale, n),
               hipLaunchKernelGGL(k_scale_constrain<double2>, grid, dim3(kRowThreads), 0, st,
                                  (double2*)C, ldb, ldk, (int)nb, (int)n_tot, scale, n));
  return QMX_OK;
}

int qmx_randomize(int dtype, void* C, int64_t ldb, int64_t ldk, int64_t nk, int64_t n_tot,
                  int64_t n_max, int64_t k_offset, const int32_t* n, const double* ke,
                  int64_t ke_ldk, int64_t seed, int64_t b_start, int64_t b_stop, void* stream) {
  QMX_CHECK_ARG(C && n && ke, "qmx_randomize: null argument");
  QMX_CHECK_ARG(nk >= 0 && nk < 65536 && n_tot >= 0 && n_tot < (1 << 30) && b_start >= 0,
                "qmx_randomize: bad sizes");
  if (nk == 0 || n_tot == 0 || b_stop <= b_start) return QMX_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  dim3 grid((unsigned)((n_tot + 255) / 256), (unsigned)nk);
  QMX_DISPATCH(dtype,
               hipLaunchKernelGGL(k_randomize<float2>, grid, dim3(256), 0, st, (float2*)C, ldb,
                                  ldk, (int)n_tot, n_max, k_offset, n, ke, ke_ldk, seed, b_start,
                                  b_stop),
               hipLaunchKernelGGL(k_randomize<double2>, grid, dim3(256), 0, st, (double2*)C, ldb,
                                  ldk, (int)n_tot, n_max, k_offset, n, ke, ke_ldk, seed, b_start,
                                  b_stop));
  return QMX_OK;
}

int qmx_apply_gradient(int dtype, const void* C, int64_t ldb, int64_t ldk, int64_t nk,
                       int64_t nb, int64_t n_tot, const void* G, int dir, double alpha, void* out,
                       int64_t out_ldb, int64_t out_ldk, int accumulate, void* stream) {
  QMX_CHECK_ARG(C && G && out, "qmx_apply_gradient: null argument");
  QMX_CHECK_ARG(nk >= 0 && nk < 65536 && nb >= 0 && nb < 65536 && n_tot >= 0 &&
                    n_tot < (1 << 30),
                "qmx_apply_gradient: bad sizes");
  QMX_CHECK_ARG(dir >= 0 && dir < 3, "qmx_apply_gradient: dir must be 0, 1 or 2");
  QMX_CHECK_ARG(ldb >= n_tot && out_ldb >= n_tot, "qmx_apply_gradient: bad strides");
  if (nk == 0 || nb == 0 || n_tot == 0) return QMX_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  dim3 grid((unsigned)((n_tot + 255) / 256), (unsigned)nb, (unsigned)nk);
  QMX_DISPATCH(dtype,
               hipLaunchKernelGGL((k_apply_gradient<float2, float>), grid, dim3(256), 0, st,
                                  (const float2*)C, ldb, ldk, (float2*)out, out_ldb, out_ldk,
                                  (int)n_tot, (const float*)G, dir, (float)alpha, accumulate),
               hipLaunchKernelGGL((k_apply_gradient<double2, double>), grid, dim3(256), 0, st,
                                  (const double2*)C, ldb, ldk, (double2*)out, out_ldb, out_ldk,
                                  (int)n_tot, (const double*)G, dir, alpha, accumulate));
  return QMX_OK;
}

int qmx_axpby(int dtype, int64_t nk, int64_t nb, int64_t n_tot, const double* a, const double* b,
              const double* c, int64_t cf_ldk, int64_t cf_ldb, const void* X, const void* Y,
              const void* Z, void* out, int64_t ldb, int64_t ldk, void* stream) {
  QMX_CHECK_ARG(nk >= 0 && nb >= 0 && n_tot >= 0, "qmx_axpby: negative size");
  if (nk == 0 || nb == 0 || n_tot == 0) return QMX_OK;
  QMX_CHECK_ARG(a && X && out, "qmx_axpby: null argument");
  QMX_CHECK_ARG(!b || Y, "qmx_axpby: coefficient b without Y");
  QMX_CHECK_ARG(ldb >= n_tot && (nk == 1 || ldk >= ldb * nb), "qmx_axpby: bad strides");
  QMX_CHECK_ARG(nk < 65536 && nb < (1ll << 31), "qmx_axpby: too many rows");
  hipStream_t st = static_cast<hipStream_t>(stream);
  dim3 grid((unsigned)nb, (unsigned)nk);
  QMX_DISPATCH(dtype,
               hipLaunchKernelGGL(k_axpby<float2>, grid, dim3(kRowThreads), 0, st, a, b, c,
                                  cf_ldk, cf_ldb, static_cast<const float2*>(X),
                                  static_cast<const float2*>(Y), static_cast<const float2*>(Z),
                                  static_cast<float2*>(out), ldb, ldk, (int)n_tot),
               hipLaunchKernelGGL(k_axpby<double2>, grid, dim3(kRowThreads), 0, st, a, b, c,
                                  cf_ldk, cf_ldb, static_cast<const double2*>(X),
                                  static_cast<const double2*>(Y), static_cast<const double2*>(Z),
                                  static_cast<double2*>(out), ldb, ldk, (int)n_tot));
  return QMX_OK;
}

int qmx_copy_blocks(const void* blocks, int64_t n, int64_t n_wg, void* stream) {
  QMX_CHECK_ARG(blocks && n >= 1 && n < (1ll << 31), "qmx_copy_blocks: bad block list");
  QMX_CHECK_ARG(n_wg >= 1 && n_wg < (1ll << 31), "qmx_copy_blocks: bad workgroup count");
  hipLaunchKernelGGL(k_copy_blocks, dim3((unsigned)n_wg), dim3(kRowThreads), 0,
                     static_cast<hipStream_t>(stream), static_cast<const CopyBlock*>(blocks),
                     (int)n);
  QMX_LAUNCHED();
  return QMX_OK;
}

}  // extern "C"
