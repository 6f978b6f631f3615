// The fused plane pass of the local apply with the plane size fixed at
// compile time, for the cubic boxes of the batched and served workloads
// (24, 30, 32, 36, 40, 48, 60 and 64 points per axis; planes up to 64 x 64).
//
// Same work as k_plane_potential (qmx_local.hip): one workgroup of 256
// threads per (z, band, k) x-y plane held in LDS; scatter the plane's active
// columns, inverse y-transform on the active x rows, inverse x-transform on
// every y column, multiply by V(x, y, z), forward x, forward y, gather back.
// Each 1D transform is two in-place Stockham passes whose radices, strides,
// butterfly counts and LDS pitch are constants here, so the butterfly
// addressing reduces to constant offsets from one base per butterfly and the
// index divisions to multiplications by constants. The runtime-general
// kernel spends about as many VALU instructions on that addressing
// (including quarter-rate 32-bit multiplies) as on the transforms.
#include "qmx_fft_core.h"
#include "qmx_plane_fixed.h"

using namespace qmx;

namespace {

constexpr int kNT = 256;
constexpr int kColPer = 4;   // active columns staged per thread (ncol <= 1024)
constexpr int kVPer = 16;    // potential values per thread (planes <= 4096 points)

// First and second radix of a length-N transform (composite radices run in
// registers, dft_any); coprime pairs take the prime-factor passes.
template <int N> struct Split;
template <> struct Split<24> { static constexpr int A = 8, B = 3; };
template <> struct Split<30> { static constexpr int A = 10, B = 3; };
template <> struct Split<32> { static constexpr int A = 8, B = 4; };
template <> struct Split<36> { static constexpr int A = 9, B = 4; };
template <> struct Split<40> { static constexpr int A = 8, B = 5; };
template <> struct Split<48> { static constexpr int A = 16, B = 3; };
template <> struct Split<60> { static constexpr int A = 15, B = 4; };
template <> struct Split<64> { static constexpr int A = 16, B = 4; };

// Thread t's butterfly j and line l over nl row lines, rows fastest:
// consecutive lanes take the same butterfly of consecutive rows, so their
// LDS addresses differ by the odd row pitch LD and a wave's 8-byte reads and
// writes fall on distinct banks (butterflies fastest would put them B slots
// apart, on banks shared by several lanes). t / nl through a float
// reciprocal: (t + 1/2) / nl is at least 1/(2 nl) >= 1/128 from an integer
// and t < 2^13, so the rounding of the product (a few ulp) never reaches the
// next integer.
__device__ __forceinline__ void row_line(int t, int nl, float rinv, int& j, int& l) {
  j = static_cast<int>((static_cast<float>(t) + 0.5f) * rinv);
  l = t - j * nl;
}

__device__ __forceinline__ float row_rinv(int nl) {
  return __builtin_amdgcn_rcpf(static_cast<float>(nl > 0 ? nl : 1));
}

// Lane map of each pass (first, second) of the row transforms (true: rows
// fastest, row_line; false: butterflies fastest) and of the column
// transforms (true: one butterfly per wave, lanes on the NY columns; false:
// lanes packed over (butterfly, column), a wave's lanes wrapping to the next
// butterfly mid-wave). gfx950's LDS serves an 8-byte access 32 lanes a
// cycle over 64 banks of 4 bytes (an LDS probe: strides of 8, 24, 40 and
// odd multiples of 8 bytes take 2 cycles a wave, 16 bytes 4, 32 bytes 8).
// At 60 points the packed column map costs 3 cycles for 2 where a wave
// wraps, and the radix-4 row pass reads 32-byte strided slots: the maps
// below cut the plane's LDS cycles by a fifth there and its time by 2.5%.
// At the other sizes the packed maps were as fast or faster when timed
// (fewer partly filled waves), so they keep them.
template <int N> struct LaneMap { static constexpr bool row0 = false, row1 = false, col0 = false, col1 = false; };
template <> struct LaneMap<60> { static constexpr bool row0 = true, row1 = false, col0 = true, col1 = true; };

// Thread t's (butterfly j, line lo offset) of a pass with NBF butterflies
// per line over nl lines (the row offset through rows[] in ROWS mode); false
// where t is past the pass's work.
template <int NBF, int NLMAX, bool ROWS, bool MAP, int LD>
__device__ __forceinline__ bool pass_item(int it, int nl, float rinv, const int* rows, int& j,
                                          int& lo) {
  if constexpr (ROWS) {
    const int t = threadIdx.x + it * kNT;
    int l;
    if constexpr (MAP) {
      row_line(t, nl, rinv, j, l);
    } else {
      l = t / NBF;
      j = t - l * NBF;
    }
    const bool on = t < NBF * nl;
    lo = on ? rows[l] * LD : 0;
    return on;
  } else if constexpr (MAP) {
    const int lane = threadIdx.x & 63;
    j = it * (kNT / 64) + (threadIdx.x >> 6);
    lo = lane;
    return lane < NLMAX && j < NBF;
  } else {
    const int t = threadIdx.x + it * kNT;
    j = t / NLMAX;
    lo = t - j * NLMAX;
    return t < NBF * nl;
  }
}

// Iterations of a pass over the block's threads.
template <int NBF, int NLMAX, bool ROWS, bool MAP>
constexpr int pass_iters() {
  return (!ROWS && MAP) ? (NBF + kNT / 64 - 1) / (kNT / 64) : (NBF * NLMAX + kNT - 1) / kNT;
}

// One in-place Stockham pass of radix R (sub-transform length NS before it)
// over nl lines of length N. ROWS: line l is row rows[l] of the plane, its
// points contiguous; otherwise line l is column l (NL columns, stride LD
// between points). Threads map to (butterfly, line) by pass_item under the
// pass's LaneMap. Values are held in registers across the barrier that
// separates the reads from the writes.
template <typename C2, int S, int N, int R, int NS, int NLMAX, bool ROWS, bool MAP, int LD>
__device__ __forceinline__ void fixed_pass(C2* buf, const C2* __restrict__ tw, int nl,
                                           const int* __restrict__ rows) {
  constexpr int STRIDE = N / R, TWSTEP = N / (NS * R);
  constexpr int NB = pass_iters<STRIDE, NLMAX, ROWS, MAP>();
  constexpr int PS = ROWS ? 1 : LD;  // distance between the points of a line
  const float rinv = ROWS ? row_rinv(nl) : 0.f;
  C2 v[NB][R];
  int out0[NB];
  bool on[NB];
#pragma unroll
  for (int it = 0; it < NB; ++it) {
    int j, lo;
    on[it] = pass_item<STRIDE, NLMAX, ROWS, MAP, LD>(it, nl, rinv, rows, j, lo);
    const int k = j % NS;
    const int in0 = j * PS + lo;
    out0[it] = ((j - k) * R + k) * PS + lo;
    if (on[it]) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[it][r] = buf[in0 + r * STRIDE * PS];
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          C2 w = tw[r * k * TWSTEP];
          if constexpr (S > 0) w.y = -w.y;
          v[it][r] = cmul(v[it][r], w);
        }
      }
      dft_any<C2, R, S>(v[it], tw, N);
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < NB; ++it) {
    if (on[it]) {
#pragma unroll
      for (int r = 0; r < R; ++r) buf[out0[it] + r * NS * PS] = v[it][r];
    }
  }
  __syncthreads();
}

// Prime-factor (Good-Thomas) form of a length N = A B transform with
// gcd(A, B) = 1, which needs no twiddles between its passes. The line is
// held permuted: space index n = (B n1 + A n2) mod N at slot B n1 + n2 (the
// rho order), frequency index k at slot B (k mod A) + (k mod B) (the gamma
// order). Either direction is then pass 1, DFT_A over slots j + B r
// (butterfly j < B), and pass 2, DFT_B over slots B j + r (butterfly
// j < A), each butterfly in place on its own slots (no barrier between a
// pass's reads and writes): gamma -> rho for the inverse transform,
// rho -> gamma for the forward one. Lines and threads as in fixed_pass.
template <int N>
struct Pfa {
  static constexpr int A = Split<N>::A, B = Split<N>::B;
  // (24 = 8 x 3 keeps the twiddled passes: its radix-8 butterflies keep
  // their internal twiddles, and the gamma index maps cost more VALU than
  // the prime-factor passes save.)
  static constexpr bool on = gcd_int(A, B) == 1 && N != 24;
  __device__ static int gamma(int k) { return (k % A) * B + k % B; }
};

template <typename C2, int S, int N, bool FIRST, int NLMAX, bool ROWS, bool MAP, int LD>
__device__ __forceinline__ void fixed_pass_pfa(C2* buf, const C2* __restrict__ tw, int nl,
                                               const int* __restrict__ rows) {
  constexpr int A = Pfa<N>::A, B = Pfa<N>::B;
  constexpr int R = FIRST ? A : B, NBF = N / R;  // radix, butterflies per line
  constexpr int STEP = FIRST ? B : 1;            // slot step within a butterfly
  constexpr int NB = pass_iters<NBF, NLMAX, ROWS, MAP>();
  constexpr int PS = ROWS ? 1 : LD;
  const float rinv = ROWS ? row_rinv(nl) : 0.f;
#pragma unroll
  for (int it = 0; it < NB; ++it) {
    int j, lo;
    if (pass_item<NBF, NLMAX, ROWS, MAP, LD>(it, nl, rinv, rows, j, lo)) {
      C2* base = buf + (FIRST ? j : j * B) * PS + lo;
      C2 v[R];
#pragma unroll
      for (int r = 0; r < R; ++r) v[r] = base[r * STEP * PS];
      dft_any<C2, R, S>(v, tw, N);
#pragma unroll
      for (int r = 0; r < R; ++r) base[r * STEP * PS] = v[r];
    }
  }
  __syncthreads();
}

template <typename C2, int S, int N, int NLMAX, bool ROWS, int LD>
__device__ __forceinline__ void fixed_fft(C2* buf, const C2* tw, int nl, const int* rows) {
  constexpr int A = Split<N>::A, B = Split<N>::B;
  static_assert(A * B == N, "radix split");
  constexpr bool M0 = ROWS ? LaneMap<N>::row0 : LaneMap<N>::col0;
  constexpr bool M1 = ROWS ? LaneMap<N>::row1 : LaneMap<N>::col1;
  static_assert(ROWS || !(M0 || M1) || NLMAX <= 64, "one butterfly per wave needs NY <= 64");
  if constexpr (Pfa<N>::on) {
    fixed_pass_pfa<C2, S, N, true, NLMAX, ROWS, M0, LD>(buf, tw, nl, rows);
    fixed_pass_pfa<C2, S, N, false, NLMAX, ROWS, M1, LD>(buf, tw, nl, rows);
  } else {
    fixed_pass<C2, S, N, A, 1, NLMAX, ROWS, M0, LD>(buf, tw, nl, rows);
    fixed_pass<C2, S, N, B, A, NLMAX, ROWS, M1, LD>(buf, tw, nl, rows);
  }
}

template <typename C2, typename R, int NX, int NY>
__global__ void __launch_bounds__(kNT)
    k_plane_fixed(C2* __restrict__ T1, const R* __restrict__ Vt, const int* __restrict__ vslot,
                  const int* __restrict__ col_ix, const int* __restrict__ col_iy,
                  const int* __restrict__ x_list, const C2* __restrict__ twx_g,
                  const C2* __restrict__ twy_g, int ncol, int nxa, int nb) {
  constexpr int LD = NY | 1, NXY = NX * NY;
  static_assert(NXY <= kNT * kVPer, "plane too large for the potential prefetch");
  extern __shared__ __align__(16) unsigned char smem_fixed[];
  const int nz = gridDim.x;
  C2* P = reinterpret_cast<C2*>(smem_fixed);
  C2* twx = P + NX * LD;
  C2* twy = twx + NX;
  int* xl = reinterpret_cast<int*>(twy + NY);
  const int z = blockIdx.x, ib = blockIdx.y, k = blockIdx.z;
  const int tid = threadIdx.x;
  C2* row = T1 + (((int64_t)k * nb + ib) * nz + z) * ncol;
  // The staged columns as separate real and imaginary parts: held in an
  // array of double2 across the zero fill, they were kept in scratch memory.
  R cvx[kColPer], cvy[kColPer];
  int co[kColPer];
#pragma unroll
  for (int q = 0; q < kColPer; ++q) {
    const int c = tid + q * kNT;
    if (c < ncol) {
      const C2 t = row[c];
      cvx[q] = t.x;
      cvy[q] = t.y;
      if constexpr (Pfa<NX>::on)  // frequency indices in the gamma order
        co[q] = Pfa<NX>::gamma(col_ix[c]) * LD + Pfa<NY>::gamma(col_iy[c]);
      else
        co[q] = col_ix[c] * LD + col_iy[c];
    }
  }
  for (int i = tid; i < NX; i += kNT) twx[i] = twx_g[i];
  for (int i = tid; i < NY; i += kNT) twy[i] = twy_g[i];
  for (int i = tid; i < nxa; i += kNT) {
    if constexpr (Pfa<NX>::on)
      xl[i] = Pfa<NX>::gamma(x_list[i]);
    else
      xl[i] = x_list[i];
  }
  for (int i = tid; i < NX * LD; i += kNT) P[i] = C2{R(0), R(0)};
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kColPer; ++q)
    if (tid + q * kNT < ncol) P[co[q]] = C2{cvx[q], cvy[q]};
  __syncthreads();
  fixed_fft<C2, +1, NY, NX, true, LD>(P, twy, nxa, xl);
  // The potential of the plane, in registers across the inverse x transform
  // (for prime-factor sizes the potential boxes come in the rho order of x
  // and y, the order of the plane's slots after the inverse transforms).
  const R* Vz = Vt + ((int64_t)vslot[k] * nz + z) * NXY;
  constexpr int VQ = (NXY + kNT - 1) / kNT;
  R vv[VQ];
#pragma unroll
  for (int q = 0; q < VQ; ++q) {
    const int i = tid + q * kNT;
    vv[q] = i < NXY ? Vz[i] : R(0);
  }
  fixed_fft<C2, +1, NX, NY, false, LD>(P, twx, NY, nullptr);
#pragma unroll
  for (int q = 0; q < VQ; ++q) {
    const int i = tid + q * kNT;
    if (i < NXY) {
      const int x = i / NY, y = i - x * NY;
      P[x * LD + y] = cscale(P[x * LD + y], vv[q]);
    }
  }
  __syncthreads();
  fixed_fft<C2, -1, NX, NY, false, LD>(P, twx, NY, nullptr);
  fixed_fft<C2, -1, NY, NX, true, LD>(P, twy, nxa, xl);
#pragma unroll
  for (int q = 0; q < kColPer; ++q)
    if (tid + q * kNT < ncol) row[tid + q * kNT] = P[co[q]];
}

template <typename C2, typename R, int N>
int launch_n(dim3 grid, size_t lds, hipStream_t st, void* T1, const void* Vt, const int* vslot,
             const int* col_ix, const int* col_iy, const int* x_list, const void* twx,
             const void* twy, int ncol, int nxa, int nb) {
  static size_t lds_set = 0;
  if (lds > lds_set) {
    QMX_HIP(hipFuncSetAttribute((const void*)k_plane_fixed<C2, R, N, N>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    lds_set = lds;
  }
  hipLaunchKernelGGL((k_plane_fixed<C2, R, N, N>), grid, dim3(kNT), lds, st,
                     static_cast<C2*>(T1), static_cast<const R*>(Vt), vslot, col_ix, col_iy,
                     x_list, static_cast<const C2*>(twx), static_cast<const C2*>(twy), ncol, nxa,
                     nb);
  QMX_LAUNCHED();
  return QMX_OK;
}

template <typename C2, typename R>
int launch_any(int n, dim3 grid, size_t lds, hipStream_t st, void* T1, const void* Vt,
               const int* vslot, const int* col_ix, const int* col_iy, const int* x_list,
               const void* twx, const void* twy, int ncol, int nxa, int nb) {
#define QMX_FIXED(N) \
  case N: return launch_n<C2, R, N>(grid, lds, st, T1, Vt, vslot, col_ix, col_iy, x_list, twx, twy, ncol, nxa, nb);
  switch (n) {
    QMX_FIXED(24)
    QMX_FIXED(30)
    QMX_FIXED(32)
    QMX_FIXED(36)
    QMX_FIXED(40)
    QMX_FIXED(48)
    QMX_FIXED(60)
    QMX_FIXED(64)
  }
#undef QMX_FIXED
  set_error("plane_fixed: no kernel for this plane size");
  return QMX_ERR_UNSUPPORTED;
}

}  // namespace

namespace qmx {

bool plane_fixed_space_order(int n, int* A, int* B) {
  int a = 0, b = 0;
  switch (n) {
#define QMX_PFA(N)                          \
  case N:                                   \
    if (Pfa<N>::on) {                       \
      a = Pfa<N>::A;                        \
      b = Pfa<N>::B;                        \
    }                                       \
    break;
    QMX_PFA(24)
    QMX_PFA(30)
    QMX_PFA(32)
    QMX_PFA(36)
    QMX_PFA(40)
    QMX_PFA(48)
    QMX_PFA(60)
    QMX_PFA(64)
#undef QMX_PFA
  }
  *A = a;
  *B = b;
  return a > 0;
}

bool plane_fixed_available(int nx, int ny, int ncol) {
  if (nx != ny || ncol > kColPer * kNT) return false;
  switch (nx) {
    case 24: case 30: case 32: case 36: case 40: case 48: case 60: case 64: return true;
  }
  return false;
}

int launch_plane_fixed(bool c128, int n, dim3 grid, size_t lds, hipStream_t st, void* T1,
                       const void* Vt, const int* vslot, const int* col_ix, const int* col_iy,
                       const int* x_list, const void* twx, const void* twy, int ncol, int nxa,
                       int nb) {
  if (c128)
    return launch_any<double2, double>(n, grid, lds, st, T1, Vt, vslot, col_ix, col_iy, x_list,
                                       twx, twy, ncol, nxa, nb);
  return launch_any<float2, float>(n, grid, lds, st, T1, Vt, vslot, col_ix, col_iy, x_list, twx,
                                   twy, ncol, nxa, nb);
}

}  // namespace qmx
