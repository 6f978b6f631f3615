// The column passes of the local apply (first and last pass of the plane
// path) with the z length fixed at compile time, for the boxes of the
// batched and served workloads (24, 30, 32, 36, 40, 48, 60 and 64 points).
//
// Same work as k_sphere_to_cols / k_cols_to_sphere (qmx_local.hip): the
// sphere points of a band scatter into their active (x, y) columns, each
// column is transformed along z, and T1 holds the columns z-major
// (T1[z][col]); the reverse pass transforms back, gathers the sphere,
// scales by 1/N and adds KE C. Here one thread owns one column: its z line
// lives in registers and is transformed there (two radix stages of
// compile-time size, no LDS round trips and no barriers inside the
// transform). LDS carries only the band's sphere points of the
// workgroup's columns, so that the sphere side is read and written in
// coalesced runs: every column's points are one cyclic interval of z (a
// sphere meets a line in an interval), recorded per (k, column) by the plan
// as (lo, hi): the points are z = lo..hi in order when lo >= 0, and
// z = 0..hi followed by z = n + lo..n - 1 when lo < 0.
#include "qmx_fft_core.h"
#include "qmx_plane_fixed.h"

using namespace qmx;

namespace {

constexpr int kCT = kColFixedThreads;  // columns (one per thread) per workgroup

// First and second radix of the in-register transform (A is composite).
template <int N> struct ZSplit;
template <> struct ZSplit<24> { static constexpr int A = 8, B = 3; };
template <> struct ZSplit<30> { static constexpr int A = 10, B = 3; };
template <> struct ZSplit<32> { static constexpr int A = 8, B = 4; };
template <> struct ZSplit<36> { static constexpr int A = 9, B = 4; };
template <> struct ZSplit<40> { static constexpr int A = 8, B = 5; };
template <> struct ZSplit<48> { static constexpr int A = 16, B = 3; };
template <> struct ZSplit<60> { static constexpr int A = 15, B = 4; };
template <> struct ZSplit<64> { static constexpr int A = 16, B = 4; };

// Transform of one line held in registers, v[n] = x[n] on entry. With
// n = B n1 + n2 and k = k1 + A k2:
//   X[k] = sum_n2 W_N^(n2 k1) W_B^(n2 k2) sum_n1 x[B n1 + n2] W_A^(n1 k1),
// B transforms of size A in place (k1 replaces n1 in slot B k1 + n2, times
// the twiddle), then A transforms of size B in place, so X[k1 + A k2] ends
// in slot B k1 + k2 (out_slot). tw is the exact N-point table; S = +1
// inverse (unscaled), -1 forward.
//
// Coprime A and B take the prime-factor form instead (no twiddles between
// the stages): the size-A transforms run over the slots (B n1 + A n2) mod N
// (n1 < A) and the size-B ones over (B k1 + A n2) mod N (n2 < B), each in
// place, so X[k] ends in slot (B (k mod A) + A (k mod B)) mod N.
template <int N>
constexpr bool zsplit_pfa() {
  return gcd_int(ZSplit<N>::A, ZSplit<N>::B) == 1;
}

template <typename C2, int N, int S>
__device__ __forceinline__ void line_fft(C2 (&v)[N], const C2* __restrict__ tw) {
  constexpr int A = ZSplit<N>::A, B = ZSplit<N>::B;
  static_assert(A * B == N, "radix split");
  if constexpr (zsplit_pfa<N>()) {
    // In place: each transform writes its outputs back to the slots it read.
#pragma unroll
    for (int n2 = 0; n2 < B; ++n2) {
      C2 a[A];
#pragma unroll
      for (int n1 = 0; n1 < A; ++n1) a[n1] = v[(B * n1 + A * n2) % N];
      dft_any<C2, A, S>(a, tw, N);
#pragma unroll
      for (int k1 = 0; k1 < A; ++k1) v[(B * k1 + A * n2) % N] = a[k1];
    }
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) {
      C2 b[B];
#pragma unroll
      for (int n2 = 0; n2 < B; ++n2) b[n2] = v[(B * k1 + A * n2) % N];
      dft_any<C2, B, S>(b, tw, N);
#pragma unroll
      for (int k2 = 0; k2 < B; ++k2) v[(B * k1 + A * k2) % N] = b[k2];
    }
    return;
  }
#pragma unroll
  for (int n2 = 0; n2 < B; ++n2) {
    C2 a[A];
#pragma unroll
    for (int n1 = 0; n1 < A; ++n1) a[n1] = v[B * n1 + n2];
    dft_any<C2, A, S>(a, tw, N);
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) {
      const int m = (n2 * k1) % N;
      if (m != 0) {
        C2 w = tw[m];
        if constexpr (S > 0) w.y = -w.y;
        a[k1] = cmul(a[k1], w);
      }
      v[B * k1 + n2] = a[k1];
    }
  }
#pragma unroll
  for (int k1 = 0; k1 < A; ++k1) {
    C2 b[B];
#pragma unroll
    for (int n2 = 0; n2 < B; ++n2) b[n2] = v[B * k1 + n2];
    dft_any<C2, B, S>(b, tw, N);
#pragma unroll
    for (int k2 = 0; k2 < B; ++k2) v[B * k1 + k2] = b[k2];
  }
}

template <int N>
__device__ constexpr int out_slot(int z) {
  if constexpr (zsplit_pfa<N>())
    return (ZSplit<N>::B * (z % ZSplit<N>::A) + ZSplit<N>::A * (z % ZSplit<N>::B)) % N;
  else
    return ZSplit<N>::B * (z % ZSplit<N>::A) + z / ZSplit<N>::A;
}

// Position of point z in its column's run, -1 when z is not a point of it.
__device__ __forceinline__ int run_offset(int z, int lo, int hi, int n) {
  if (lo >= 0) return (z >= lo && z <= hi) ? z - lo : -1;
  if (z <= hi) return z;
  return z >= n + lo ? hi + 1 + (z - (n + lo)) : -1;
}

__device__ __forceinline__ void unpack_zrange(int zr, int& lo, int& hi) {
  lo = (int)(short)(zr & 0xFFFF);
  hi = zr >> 16;
}

template <typename C2, typename R, int NZ>
__global__ void __launch_bounds__(kCT)
    k_sphere_to_cols_fixed(const C2* __restrict__ C, int64_t ldb, int64_t ldk,
                           C2* __restrict__ T1, const int* __restrict__ cstart,
                           const int* __restrict__ zrange, const C2* __restrict__ tw_g, int ncol,
                           int nb, int b0) {
  extern __shared__ __align__(16) unsigned char smem_cols[];
  C2* tw = reinterpret_cast<C2*>(smem_cols);
  C2* L = tw + NZ;  // the sphere points of this workgroup's columns
  const int tid = threadIdx.x;
  const int col0 = blockIdx.x * kCT, ib = blockIdx.y, k = blockIdx.z;
  const int mcols = min(kCT, ncol - col0);
  const int* cs = cstart + (int64_t)k * (ncol + 1);
  const int gs = cs[col0], ge = cs[col0 + mcols];
  const C2* Cb = C + k * ldk + (int64_t)(b0 + ib) * ldb;
  for (int i = tid; i < NZ; i += kCT) tw[i] = tw_g[i];
  for (int g = gs + tid; g < ge; g += kCT) L[g - gs] = Cb[g];
  const bool live = tid < mcols;
  int r0 = 0, lo = 1, hi = 0;  // an empty run unless live
  if (live) {
    r0 = cs[col0 + tid] - gs;
    unpack_zrange(zrange[(int64_t)k * ncol + col0 + tid], lo, hi);
  }
  __syncthreads();
  C2 v[NZ];
#pragma unroll
  for (int z = 0; z < NZ; ++z) {
    const int o = run_offset(z, lo, hi, NZ);
    v[z] = o >= 0 ? L[r0 + o] : C2{R(0), R(0)};
  }
  line_fft<C2, NZ, +1>(v, tw);
  if (!live) return;
  C2* col = T1 + ((int64_t)k * nb + ib) * (int64_t)ncol * NZ + col0 + tid;
#pragma unroll
  for (int z = 0; z < NZ; ++z) col[(int64_t)z * ncol] = v[out_slot<NZ>(z)];
}

template <typename C2, typename R, int NZ>
__global__ void __launch_bounds__(kCT)
    k_cols_to_sphere_fixed(const C2* __restrict__ T1, const C2* C, int64_t ldb, int64_t ldk,
                           C2* out, int64_t oldb, int64_t oldk, const R* __restrict__ ke,
                           int64_t ke_ldk, const int* __restrict__ cstart,
                           const int* __restrict__ zrange, const int* __restrict__ nvec,
                           const C2* __restrict__ tw_g, int ncol, int nb, int n_tot, int b0,
                           R scale, int accumulate) {
  extern __shared__ __align__(16) unsigned char smem_cols[];
  C2* tw = reinterpret_cast<C2*>(smem_cols);
  C2* L = tw + NZ;
  const int tid = threadIdx.x;
  const int col0 = blockIdx.x * kCT, ib = blockIdx.y, k = blockIdx.z;
  const int mcols = min(kCT, ncol - col0);
  const bool live = tid < mcols;
  const C2* col = T1 + ((int64_t)k * nb + ib) * (int64_t)ncol * NZ + col0 + (live ? tid : 0);
  C2 v[NZ];
#pragma unroll
  for (int z = 0; z < NZ; ++z) v[z] = live ? col[(int64_t)z * ncol] : C2{R(0), R(0)};
  for (int i = tid; i < NZ; i += kCT) tw[i] = tw_g[i];
  const int* cs = cstart + (int64_t)k * (ncol + 1);
  const int gs = cs[col0], ge = cs[col0 + mcols];
  int r0 = 0, lo = 1, hi = 0;
  if (live) {
    r0 = cs[col0 + tid] - gs;
    unpack_zrange(zrange[(int64_t)k * ncol + col0 + tid], lo, hi);
  }
  __syncthreads();
  line_fft<C2, NZ, -1>(v, tw);
#pragma unroll
  for (int z = 0; z < NZ; ++z) {
    const int o = run_offset(z, lo, hi, NZ);
    if (o >= 0) L[r0 + o] = v[out_slot<NZ>(z)];
  }
  __syncthreads();
  const int64_t band = b0 + ib;
  const C2* Cb = C + k * ldk + band * ldb;
  C2* ob = out + k * oldk + band * oldb;
  const R* keb = ke ? ke + k * ke_ldk : nullptr;
  for (int g = gs + tid; g < ge; g += kCT) {
    C2 val = cscale(L[g - gs], scale);
    if (keb) {
      const C2 c = Cb[g];
      const R e = keb[g];
      val.x += e * c.x;
      val.y += e * c.y;
    }
    if (accumulate) val = cadd(val, ob[g]);
    ob[g] = val;
  }
  if (blockIdx.x == 0) {
    for (int g = nvec[k] + tid; g < n_tot; g += kCT) ob[g] = C2{R(0), R(0)};
  }
}

// One static per kernel (LAST: the last pass), so each kernel gets its own
// attribute call.
template <typename C2, typename R, int NZ, bool LAST>
int set_lds(const void* fn, size_t lds) {
  static size_t done = 0;
  if (lds > done) {
    QMX_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    done = lds;
  }
  return QMX_OK;
}

template <typename C2, typename R, int NZ>
int launch_a(const ColFixedArgs& a) {
  const size_t lds = (NZ + (size_t)a.run_max) * sizeof(C2);
  const void* fn = reinterpret_cast<const void*>(&k_sphere_to_cols_fixed<C2, R, NZ>);
  int rc = set_lds<C2, R, NZ, false>(fn, lds);
  if (rc) return rc;
  hipLaunchKernelGGL((k_sphere_to_cols_fixed<C2, R, NZ>), a.grid, dim3(kCT), lds, a.st,
                     static_cast<const C2*>(a.C), a.ldb, a.ldk, static_cast<C2*>(a.T1), a.cstart,
                     a.zrange, static_cast<const C2*>(a.tw), a.ncol, a.nb, a.b0);
  QMX_LAUNCHED();
  return QMX_OK;
}

template <typename C2, typename R, int NZ>
int launch_e(const ColFixedArgs& a) {
  const size_t lds = (NZ + (size_t)a.run_max) * sizeof(C2);
  const void* fn = reinterpret_cast<const void*>(&k_cols_to_sphere_fixed<C2, R, NZ>);
  int rc = set_lds<C2, R, NZ, true>(fn, lds);
  if (rc) return rc;
  hipLaunchKernelGGL((k_cols_to_sphere_fixed<C2, R, NZ>), a.grid, dim3(kCT), lds, a.st,
                     static_cast<const C2*>(a.T1), static_cast<const C2*>(a.C), a.ldb, a.ldk,
                     static_cast<C2*>(a.out), a.oldb, a.oldk, static_cast<const R*>(a.ke),
                     a.ke_ldk, a.cstart, a.zrange, a.nvec, static_cast<const C2*>(a.tw), a.ncol,
                     a.nb, a.n_tot, a.b0, (R)a.scale, a.accumulate);
  QMX_LAUNCHED();
  return QMX_OK;
}

template <typename C2, typename R>
int launch_any(bool to_sphere, int nz, const ColFixedArgs& a) {
#define QMX_COLS(N) \
  case N: return to_sphere ? launch_e<C2, R, N>(a) : launch_a<C2, R, N>(a);
  switch (nz) {
    QMX_COLS(24)
    QMX_COLS(30)
    QMX_COLS(32)
    QMX_COLS(36)
    QMX_COLS(40)
    QMX_COLS(48)
    QMX_COLS(60)
    QMX_COLS(64)
  }
#undef QMX_COLS
  set_error("cols_fixed: no kernel for this z length");
  return QMX_ERR_UNSUPPORTED;
}

}  // namespace

namespace qmx {

// Complex64 only: a complex128 line of 24 or more points with its transform
// temporaries needs the whole register file of a lane (one wave per SIMD,
// spills beyond 48 points), so complex128 keeps the LDS-staged kernels.
bool cols_fixed_available(int nz, bool c128) {
  if (c128) return false;
  switch (nz) {
    case 24: case 30: case 32: case 36: case 40: case 48: case 60: case 64: return true;
  }
  return false;
}

int launch_cols_fixed(bool c128, bool to_sphere, int nz, const ColFixedArgs& a) {
  if (c128) {
    set_error("cols_fixed: complex64 only");
    return QMX_ERR_UNSUPPORTED;
  }
  return launch_any<float2, float>(to_sphere, nz, a);
}

}  // namespace qmx
