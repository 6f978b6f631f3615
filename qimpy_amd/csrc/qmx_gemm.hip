// Batched overlap product over the basis axis on the matrix cores:
//   out[b][m][n] = (accumulate ? out : 0) + sum_g conj(A[b][m][g]) B[b][n][g]
// the shape of every overlap, Rayleigh-Ritz and projector product of the
// band solve (M, N <= a few hundred bands, G = 10^3..10^5 plane waves).
//
// The basis axis is split into chunks so that the grid fills the chip even
// for one small matrix; a workgroup computes one 64 x 64 output tile of one
// batch over one chunk with 16x16x4 MFMAs (FP32 for complex64 inputs, FP64
// for complex128), staging 16-wide slabs of A and B through LDS (double
// buffered, coalesced 128-byte row segments, split into real and imaginary
// planes). Complex products take four real MFMAs:
//   Re += Ar Br + Ai Bi,   Im += Ar Bi - Ai Br.
// Chunk partials go to the workspace and a second kernel sums them in FP64
// into the complex128 result, so complex64 inputs are accumulated in FP32
// only within a chunk.
#include <algorithm>

#include "qmx_common.h"

namespace qmx {
void prof_begin(hipStream_t st, hipEvent_t* a);
void prof_end(hipStream_t st, hipEvent_t a, int id, double bytes);
}  // namespace qmx

typedef float floatx4_t __attribute__((ext_vector_type(4)));
typedef double doublex4_t __attribute__((ext_vector_type(4)));

using namespace qmx;

namespace {

constexpr int kTile = 64;      // output tile (M and N)
constexpr int kSlab = 16;      // basis-axis slab staged per step
constexpr int kDotThreads = 256;
constexpr int kMinChunk = 256;  // smallest chunk of the basis axis

template <typename T>
struct Mfma;

// Operands: lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16]. Result
// register r of lane l holds row row(l / 16, r), column l % 16.
template <>
struct Mfma<float> {
  using Acc = floatx4_t;
  __device__ static int row(int lk, int r) { return 4 * lk + r; }
  __device__ static Acc zero() { return Acc{0.f, 0.f, 0.f, 0.f}; }
  __device__ static Acc mac(float a, float b, Acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

template <>
struct Mfma<double> {
  using Acc = doublex4_t;
  __device__ static int row(int lk, int r) { return lk + 4 * r; }
  __device__ static Acc zero() { return Acc{0.0, 0.0, 0.0, 0.0}; }
  __device__ static Acc mac(double a, double b, Acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};

// One workgroup: tile (tm, tn) of batch b over basis chunk `chunk`.
// Waves in a 2 x 2 arrangement, each a 32 x 32 sub-tile = 2 x 2 MFMA tiles.
template <typename T>
__global__ void __launch_bounds__(kDotThreads)
    k_dot_partial(const typename Cplx<T>::T* __restrict__ A, int64_t lda, int64_t sa,
                  const typename Cplx<T>::T* __restrict__ B, int64_t ldb, int64_t sb, int M,
                  int N, int64_t G, int64_t chunk_len, int mtiles, int ntiles, int nchunks,
                  typename Cplx<T>::T* __restrict__ part) {
  using C2 = typename Cplx<T>::T;
  using MF = Mfma<T>;
  constexpr int LD = kSlab + 1;
  __shared__ T sAr[2][kTile * LD], sAi[2][kTile * LD], sBr[2][kTile * LD], sBi[2][kTile * LD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int bt = blockIdx.x;
  const int tn = bt % ntiles;
  bt /= ntiles;
  const int tm = bt % mtiles;
  const int b = bt / mtiles;
  const int chunk = blockIdx.y;
  const int m0 = tm * kTile, n0 = tn * kTile;
  const int64_t g_begin = (int64_t)chunk * chunk_len;
  const int64_t g_end = std::min<int64_t>(G, g_begin + chunk_len);
  const C2* Ab = A + (int64_t)b * sa;
  const C2* Bb = B + (int64_t)b * sb;
  // Staging: thread -> (row = tid / 4, 4 consecutive g at (tid % 4) * 4).
  const int srow = tid >> 2, sg = (tid & 3) * 4;
  C2 ra[4], rb[4];
  auto load = [&](int64_t g0) {
    const int64_t g = g0 + sg;
    const int am = m0 + srow, bn = n0 + srow;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const bool okg = g + q < g_end;
      ra[q] = (am < M && okg) ? Ab[(int64_t)am * lda + g + q] : C2{T(0), T(0)};
      rb[q] = (bn < N && okg) ? Bb[(int64_t)bn * ldb + g + q] : C2{T(0), T(0)};
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      sAr[buf][srow * LD + sg + q] = ra[q].x;
      sAi[buf][srow * LD + sg + q] = ra[q].y;
      sBr[buf][srow * LD + sg + q] = rb[q].x;
      sBi[buf][srow * LD + sg + q] = rb[q].y;
    }
  };
  typename MF::Acc re[2][2], im[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      re[i][j] = MF::zero();
      im[i][j] = MF::zero();
    }
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;  // wave sub-tile
  const int li = lane & 15, lk = lane >> 4;               // MFMA operand lane layout
  int buf = 0;
  if (g_begin < g_end) {
    load(g_begin);
    store(0);
  }
  __syncthreads();
  for (int64_t g0 = g_begin; g0 < g_end; g0 += kSlab) {
    const bool more = g0 + kSlab < g_end;
    if (more) load(g0 + kSlab);  // next slab in flight during the MFMAs
#pragma unroll
    for (int kk = 0; kk < kSlab; kk += 4) {
      T ar[2], ai[2], br[2], bi[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int r = (wm + 16 * i + li) * LD + kk + lk;
        ar[i] = sAr[buf][r];
        ai[i] = sAi[buf][r];
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int r = (wn + 16 * j + li) * LD + kk + lk;
        br[j] = sBr[buf][r];
        bi[j] = sBi[buf][r];
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          re[i][j] = MF::mac(ar[i], br[j], re[i][j]);
          re[i][j] = MF::mac(ai[i], bi[j], re[i][j]);
          im[i][j] = MF::mac(ar[i], bi[j], im[i][j]);
          im[i][j] = MF::mac(-ai[i], br[j], im[i][j]);
        }
    }
    if (more) {
      store(buf ^ 1);
      __syncthreads();
      buf ^= 1;
    }
  }
  // Result rows of each 16 x 16 MFMA tile as laid out by MF::row.
  C2* pb = part + ((int64_t)b * nchunks + chunk) * M * N;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + 16 * i + MF::row(lk, r), n = n0 + wn + 16 * j + li;
        if (m < M && n < N) pb[(int64_t)m * N + n] = C2{re[i][j][r], im[i][j][r]};
      }
}

// out[b][m][n] (complex128, row stride ldc, batch stride sc) = sum over chunks.
template <typename T>
__global__ void __launch_bounds__(256)
    k_dot_reduce(const typename Cplx<T>::T* __restrict__ part, int64_t batch, int M, int N,
                 int nchunks, double2* __restrict__ out, int64_t ldc, int64_t sc,
                 int accumulate) {
  const int64_t MN = (int64_t)M * N;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= batch * MN) return;
  const int64_t b = i / MN, mn = i - b * MN;
  const int m = (int)(mn / N), n = (int)(mn - (int64_t)m * N);
  double sr = 0.0, si = 0.0;
  const typename Cplx<T>::T* p = part + b * nchunks * MN + mn;
  for (int c = 0; c < nchunks; ++c) {
    sr += (double)p[c * MN].x;
    si += (double)p[c * MN].y;
  }
  double2* o = out + b * sc + (int64_t)m * ldc + n;
  if (accumulate) {
    sr += o->x;
    si += o->y;
  }
  *o = double2{sr, si};
}

struct DotPlan {
  int mtiles, ntiles, nchunks;
  int64_t chunk_len;
};

DotPlan dot_plan(int64_t batch, int64_t M, int64_t N, int64_t G) {
  DotPlan p;
  p.mtiles = (int)((M + kTile - 1) / kTile);
  p.ntiles = (int)((N + kTile - 1) / kTile);
  const int64_t tiles = std::max<int64_t>(1, batch * p.mtiles * p.ntiles);
  const int64_t want = std::max<int64_t>(1, (2048 + tiles - 1) / tiles);
  const int64_t most = std::max<int64_t>(1, (G + kMinChunk - 1) / kMinChunk);
  int64_t nch = std::min<int64_t>(std::min<int64_t>(want, most), 65535);
  int64_t len = (G + nch - 1) / nch;
  len = (len + kSlab - 1) / kSlab * kSlab;
  len = std::max<int64_t>(len, kSlab);
  p.chunk_len = len;
  p.nchunks = (int)std::max<int64_t>(1, (G + len - 1) / len);
  return p;
}

template <typename T>
int dot_impl(const void* A, int64_t lda, int64_t sa, const void* B, int64_t ldb, int64_t sb,
             int64_t batch, int64_t M, int64_t N, int64_t G, void* out, int64_t ldc, int64_t sc,
             int accumulate, void* ws, int64_t ws_bytes, hipStream_t st) {
  using C2 = typename Cplx<T>::T;
  const DotPlan p = dot_plan(batch, M, N, G);
  const int64_t need = batch * p.nchunks * M * N * (int64_t)sizeof(C2);
  QMX_CHECK_ARG(ws && ws_bytes >= need, "qmx_dot: workspace too small");
  QMX_CHECK_ARG(batch * p.mtiles * p.ntiles < (1ll << 31), "qmx_dot: too many tiles");
  hipEvent_t ev = nullptr;
  prof_begin(st, &ev);
  dim3 grid((unsigned)(batch * p.mtiles * p.ntiles), (unsigned)p.nchunks);
  hipLaunchKernelGGL((k_dot_partial<T>), grid, dim3(kDotThreads), 0, st,
                     static_cast<const C2*>(A), lda, sa, static_cast<const C2*>(B), ldb, sb,
                     (int)M, (int)N, G, p.chunk_len, p.mtiles, p.ntiles, p.nchunks,
                     static_cast<C2*>(ws));
  QMX_LAUNCHED();
  const int64_t total = batch * M * N;
  hipLaunchKernelGGL((k_dot_reduce<T>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st,
                     static_cast<const C2*>(ws), batch, (int)M, (int)N, p.nchunks,
                     static_cast<double2*>(out), ldc, sc, accumulate);
  QMX_LAUNCHED();
  prof_end(st, ev, QMX_KERNEL_GEMM + (sizeof(T) == 8 ? QMX_KERNEL_C128 : 0), (double)batch * (M + N) * G * sizeof(C2));
  return QMX_OK;
}

}  // namespace

extern "C" {

int64_t qmx_dot_workspace_bytes(int dtype, int64_t batch, int64_t M, int64_t N, int64_t G) {
  if ((dtype != QMX_C64 && dtype != QMX_C128) || batch < 0 || M < 0 || N < 0 || G < 0) return -1;
  const DotPlan p = dot_plan(batch, M, N, G);
  return std::max<int64_t>(1, batch * p.nchunks * M * N * (dtype == QMX_C64 ? 8 : 16));
}

int qmx_dot(int dtype, const void* A, int64_t lda, int64_t sa, const void* B, int64_t ldb,
            int64_t sb, int64_t batch, int64_t M, int64_t N, int64_t G, void* out, int64_t ldc,
            int64_t sc, int accumulate, void* workspace, int64_t workspace_bytes, void* stream) {
  QMX_CHECK_ARG(out, "qmx_dot: null output");
  QMX_CHECK_ARG(batch >= 0 && M >= 0 && N >= 0 && G >= 0, "qmx_dot: negative size");
  QMX_CHECK_ARG(M < (1 << 20) && N < (1 << 20), "qmx_dot: M, N too large");
  if (batch == 0 || M == 0 || N == 0) return QMX_OK;
  QMX_CHECK_ARG(A && B, "qmx_dot: null operand");
  QMX_CHECK_ARG(lda >= G && ldb >= G && ldc >= N, "qmx_dot: leading dimension too small");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == QMX_C64)
    return dot_impl<float>(A, lda, sa, B, ldb, sb, batch, M, N, G, out, ldc, sc, accumulate,
                           workspace, workspace_bytes, st);
  if (dtype == QMX_C128)
    return dot_impl<double>(A, lda, sa, B, ldb, sb, batch, M, N, G, out, ldc, sc, accumulate,
                            workspace, workspace_bytes, st);
  set_error("qmx_dot: dtype must be QMX_C64 or QMX_C128");
  return QMX_ERR_ARG;
}

}  // extern "C"
