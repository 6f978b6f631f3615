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
// planes). Complex products take three real MFMAs (Gauss's form of
// conj(a) b):
//   p1 += Ar Br,  p2 += Ai Bi,  p3 += (Ar - Ai)(Br + Bi);
//   Re = p1 + p2,  Im = p3 - p1 + p2,
// which adds one rounding of the operand sums and doubles the error bound
// of the imaginary part against the four-product form.
// Chunk partials go to the workspace and a second kernel sums them in FP64
// into the complex128 result, so complex64 inputs are accumulated in FP32
// only within a chunk.
#include <algorithm>
#include <mutex>
#include <string>
#include <type_traits>

#include "qmx_common.h"
#include "qmx_rsrc.h"
#include "qmx_tune.h"

namespace qmx {
void prof_begin(hipStream_t st, hipEvent_t* a);
void prof_end(hipStream_t st, hipEvent_t a, int id, double bytes);
void prof_end_flops(hipStream_t st, hipEvent_t a, int id, double bytes, double flops);
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
  typename MF::Acc p1[2][2], p2[2][2], p3[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      p1[i][j] = MF::zero();
      p2[i][j] = MF::zero();
      p3[i][j] = MF::zero();
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
      T ad[2], bs[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        ad[i] = ar[i] - ai[i];
        bs[i] = br[i] + bi[i];
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          p1[i][j] = MF::mac(ar[i], br[j], p1[i][j]);
          p2[i][j] = MF::mac(ai[i], bi[j], p2[i][j]);
          p3[i][j] = MF::mac(ad[i], bs[j], p3[i][j]);
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
        if (m < M && n < N)
          pb[(int64_t)m * N + n] = C2{p1[i][j][r] + p2[i][j][r],
                                      p3[i][j][r] - p1[i][j][r] + p2[i][j][r]};
      }
}

// Output blocks fitted to the matrix: a workgroup of eight waves computes a
// (16 bm16) x (16 bn16) block of one batch over one basis chunk, so the
// shapes of the band solve (144 x 72, 72 x 72, 64 x 32, 32 x 32) run without
// the padding of a fixed 64 x 64 tile. The block's 16 x 16 MFMA tiles are
// dealt to the waves round-robin (at most kMaxT each). One LDS slab holds the
// block's A rows followed by its B rows (real and imaginary planes, kSlab
// basis points per row); the next slab is loaded into registers while the
// MFMAs run. The four MFMA steps of a slab give lane (i, k) the basis points
// 4k .. 4k + 3 of its row, so each operand of a tile is one 16-byte LDS read
// per slab (two for FP64); the order of the basis sum does not matter.
constexpr int kBlkThreads = 512;
constexpr int kBlkRows16 = 14;  // bm16 + bn16 at most (224 staged rows)
constexpr int kBlkTiles = 48;   // bm16 * bn16 at most
// complex128 blocks: at most 32 tiles (four per wave); FP64 accumulators of
// six tiles a wave do not fit the registers of two waves per SIMD.
constexpr int kBlkTiles64 = 32;
constexpr int blk_tiles_max(int dtype) { return dtype == QMX_C128 ? kBlkTiles64 : kBlkTiles; }
constexpr int kBlkQuads = (kBlkRows16 * 16 * (kSlab / 4) + kBlkThreads - 1) / kBlkThreads;

// LDS row stride in elements: 16-byte aligned rows whose 16 lanes of one
// read start on distinct groups of four banks.
template <typename T>
constexpr int blk_ld() { return sizeof(T) == 4 ? 20 : 18; }

template <typename T>
size_t blk_lds_bytes(int rows16) { return (size_t)rows16 * 16 * blk_ld<T>() * 2 * sizeof(T); }

// Four consecutive LDS elements (one ds_read_b128 for FP32, two for FP64).
typedef float lds_f4 __attribute__((ext_vector_type(4)));
typedef double lds_d2 __attribute__((ext_vector_type(2)));
struct Quad32 { float v[4]; };
struct Quad64 { double v[4]; };
__device__ inline Quad32 lds_quad(const float* p) {
  const lds_f4 x = *reinterpret_cast<const lds_f4*>(p);
  return Quad32{{x[0], x[1], x[2], x[3]}};
}
__device__ inline Quad64 lds_quad(const double* p) {
  const lds_d2 x = reinterpret_cast<const lds_d2*>(p)[0], y = reinterpret_cast<const lds_d2*>(p)[1];
  return Quad64{{x[0], x[1], y[0], y[1]}};
}
__device__ inline void lds_store4(float* p, float a, float b, float c, float d) {
  *reinterpret_cast<lds_f4*>(p) = lds_f4{a, b, c, d};
}
__device__ inline void lds_store4(double* p, double a, double b, double c, double d) {
  reinterpret_cast<lds_d2*>(p)[0] = lds_d2{a, b};
  reinterpret_cast<lds_d2*>(p)[1] = lds_d2{c, d};
}

template <typename T, int kMaxT>
__global__ void __launch_bounds__(kBlkThreads)
    k_dot_block(const typename Cplx<T>::T* __restrict__ A, int64_t lda, int64_t sa,
                const typename Cplx<T>::T* __restrict__ B, int64_t ldb, int64_t sb, int M, int N,
                int64_t G, int64_t chunk_len, int bm16, int bn16, int mblocks, int nblocks,
                int nchunks, typename Cplx<T>::T* __restrict__ part) {
  using C2 = typename Cplx<T>::T;
  using MF = Mfma<T>;
  constexpr int LD = blk_ld<T>();
  constexpr int kWaves = kBlkThreads / 64;
  extern __shared__ __align__(16) unsigned char blk_smem[];
  const int rowsA = bm16 * 16, rows = rowsA + bn16 * 16, nquads = rows * (kSlab / 4);
  T* sR = reinterpret_cast<T*>(blk_smem);
  T* sI = sR + rows * LD;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int bt = blockIdx.x;
  const int bj = bt % nblocks;
  bt /= nblocks;
  const int bi = bt % mblocks;
  const int b = bt / mblocks;
  const int chunk = blockIdx.y;
  const int m0 = bi * bm16 * 16, n0 = bj * bn16 * 16;
  const int64_t g_begin = (int64_t)chunk * chunk_len;
  const int64_t g_end = std::min<int64_t>(G, g_begin + chunk_len);
  const C2* Ab = A + (int64_t)b * sa;
  const C2* Bb = B + (int64_t)b * sb;
  // Staging: quad q -> (staged row q / 4, 4 consecutive g at (q % 4) * 4).
  // Loads go through buffer resources (qmx_rsrc.h) whose base is basis
  // point g0 of A or of B: a wave stages 16 consecutive rows and rowsA is a
  // multiple of 16, so the matrix is uniform per wave. A resource ends at
  // its matrix's last point, rows past M or N take the out-of-range
  // offset, and points past the chunk (read from the next chunk or row, or
  // zero) are masked when stored, so a slab's loads are issued back to back
  // and waited for after the MFMAs of the previous slab.
  int goff[kBlkQuads], sofs[kBlkQuads], gq[kBlkQuads];
  bool fromA[kBlkQuads], staged[kBlkQuads];
  const int64_t endA = ((int64_t)(M - 1) * lda + G) * (int64_t)sizeof(C2);
  const int64_t endB = ((int64_t)(N - 1) * ldb + G) * (int64_t)sizeof(C2);
#pragma unroll
  for (int u = 0; u < kBlkQuads; ++u) {
    const int q = tid + u * kBlkThreads;
    const int row = q >> 2;
    gq[u] = (q & 3) * 4;
    sofs[u] = row * LD + gq[u];
    staged[u] = __builtin_amdgcn_readfirstlane(q < nquads);
    fromA[u] = __builtin_amdgcn_readfirstlane(row < rowsA);
    goff[u] = kOutOfRange;
    if (fromA[u]) {
      if (m0 + row < M) goff[u] = (int)(((int64_t)(m0 + row) * lda + gq[u]) * (int64_t)sizeof(C2));
    } else if (n0 + row - rowsA < N) {
      goff[u] = (int)(((int64_t)(n0 + row - rowsA) * ldb + gq[u]) * (int64_t)sizeof(C2));
    }
  }
  C2 reg[kBlkQuads][4];
  auto load = [&](int64_t g0) {
#pragma unroll
    for (int u = 0; u < kBlkQuads; ++u) {
      if (staged[u]) {
        const int64_t skip = g0 * (int64_t)sizeof(C2);
        const auto r = fromA[u] ? make_rsrc(Ab + g0, endA - skip) : make_rsrc(Bb + g0, endB - skip);
        rsrc_load4(r, goff[u], reg[u]);
      }
    }
  };
  auto store = [&](int64_t g0) {
    const bool whole = g0 + kSlab <= g_end;  // no point of the slab past the chunk
#pragma unroll
    for (int u = 0; u < kBlkQuads; ++u) {
      if (staged[u]) {
        T r[4], i[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const bool ok = whole || g0 + gq[u] + e < g_end;
          r[e] = ok ? reg[u][e].x : T(0);
          i[e] = ok ? reg[u][e].y : T(0);
        }
        lds_store4(sR + sofs[u], r[0], r[1], r[2], r[3]);
        lds_store4(sI + sofs[u], i[0], i[1], i[2], i[3]);
      }
    }
  };
  // This wave's tiles: t = wave + kWaves s. A slot past the block's last
  // tile (the host picks kMaxT = ceil(tiles / kWaves), so only a wave's last
  // slot can be one) points at that tile, skips its MFMAs and is not stored.
  const int ntiles = bm16 * bn16;
  const int li = lane & 15, lk = lane >> 4;
  int offA[kMaxT], offB[kMaxT];
  typename MF::Acc p1[kMaxT], p2[kMaxT], p3[kMaxT];
#pragma unroll
  for (int s = 0; s < kMaxT; ++s) {
    const int t = std::min(wave + kWaves * s, ntiles - 1);
    const int ti = t / bn16, tj = t - ti * bn16;
    offA[s] = (16 * ti + li) * LD + 4 * lk;
    offB[s] = (rowsA + 16 * tj + li) * LD + 4 * lk;
    p1[s] = MF::zero();
    p2[s] = MF::zero();
    p3[s] = MF::zero();
  }
  const bool last_live = wave + kWaves * (kMaxT - 1) < ntiles;  // uniform per wave
  if (g_begin < g_end) load(g_begin);
  for (int64_t g0 = g_begin; g0 < g_end; g0 += kSlab) {
    store(g0);
    // The next slab's loads go out before the barrier, which keeps the
    // compiler from sinking them below the MFMAs they should overlap
    // (clamped past the end of the chunk).
    load(g0 + kSlab);
    __syncthreads();
    // Operands of tile s + 1 are read from LDS before the MFMAs of tile s.
    auto ar = lds_quad(sR + offA[0]), ai = lds_quad(sI + offA[0]);
    auto br = lds_quad(sR + offB[0]), bi = lds_quad(sI + offB[0]);
#pragma unroll
    for (int s = 0; s < kMaxT; ++s) {
      const int sn = s + 1 < kMaxT ? s + 1 : s;
      const auto ar1 = lds_quad(sR + offA[sn]), ai1 = lds_quad(sI + offA[sn]);
      const auto br1 = lds_quad(sR + offB[sn]), bi1 = lds_quad(sI + offB[sn]);
      if (s + 1 < kMaxT || last_live) {  // a repeated last slot skips its MFMAs
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          p1[s] = MF::mac(ar.v[k], br.v[k], p1[s]);
          p2[s] = MF::mac(ai.v[k], bi.v[k], p2[s]);
          p3[s] = MF::mac(ar.v[k] - ai.v[k], br.v[k] + bi.v[k], p3[s]);
        }
      }
      ar = ar1;
      ai = ai1;
      br = br1;
      bi = bi1;
    }
    __syncthreads();
  }
  C2* pb = part + ((int64_t)b * nchunks + chunk) * M * N;
#pragma unroll
  for (int s = 0; s < kMaxT; ++s) {
    const int t = wave + kWaves * s;
    if (t < ntiles) {
      const int ti = t / bn16, tj = t - ti * bn16;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + 16 * ti + MF::row(lk, r), n = n0 + 16 * tj + li;
        if (m < M && n < N)
          pb[(int64_t)m * N + n] = C2{p1[s][r] + p2[s][r], p3[s][r] - p1[s][r] + p2[s][r]};
      }
    }
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
  bool block;          // k_dot_block (fitted blocks) or k_dot_partial (64 x 64 tiles)
  int mtiles, ntiles;  // output blocks or tiles along M and N
  int bm16, bn16;      // block size in 16-row MFMA tiles (block kernel)
  int nchunks;
  int64_t chunk_len;
};

constexpr int kBlkTarget = 1024;   // workgroups the block kernel aims for
constexpr int kTileTarget = 2048;  // workgroups the 64 x 64 kernel aims for

// Modelled cost of an output tiling, in units of one 16-row slab load per
// basis point: each 16 x 16 complex MFMA tile costs about 10 of those on the
// matrix cores (3 real MFMAs against the HBM share of one CU).
inline int64_t tiling_cost(int64_t tiles16, int64_t rows16) { return 10 * tiles16 + rows16; }

// Chunks of the basis axis: nch requested, rounded to whole slabs.
void dot_chunks(DotPlan& p, int64_t G, int64_t nch) {
  nch = std::max<int64_t>(1, std::min<int64_t>(nch, 65535));
  int64_t len = (G + nch - 1) / nch;
  len = (len + kSlab - 1) / kSlab * kSlab;
  len = std::max<int64_t>(len, kSlab);
  p.chunk_len = len;
  p.nchunks = (int)std::max<int64_t>(1, (G + len - 1) / len);
}

// A tuned choice (kernel, bm16, bn16, chunks) for this shape, if one is set
// and fits it.
bool dot_tuned(DotPlan& p, int dtype, int64_t batch, int64_t M, int64_t N, int64_t G) {
  const int64_t key[5] = {dtype, tune_b_bucket(batch), M, N, tune_g_bucket(G)};
  int64_t ch[4];
  if (!tune_lookup(QMX_TUNE_DOT, key, 5, ch, 4)) return false;
  const int64_t M16 = (M + 15) / 16, N16 = (N + 15) / 16;
  if (ch[0] == 0) {
    const int64_t bm = ch[1], bn = ch[2];
    if (bm < 1 || bn < 1 || bm > M16 || bn > N16 || bm + bn > kBlkRows16 ||
        bm * bn > blk_tiles_max(dtype))
      return false;
    p.block = true;
    p.bm16 = (int)bm;
    p.bn16 = (int)bn;
    p.mtiles = (int)((M16 + bm - 1) / bm);
    p.ntiles = (int)((N16 + bn - 1) / bn);
  } else if (ch[0] == 1) {
    p.block = false;
    p.mtiles = (int)((M + kTile - 1) / kTile);
    p.ntiles = (int)((N + kTile - 1) / kTile);
  } else {
    return false;
  }
  dot_chunks(p, G, ch[3]);
  return true;
}

DotPlan dot_plan(int dtype, int64_t batch, int64_t M, int64_t N, int64_t G) {
  DotPlan p{};
  if (dot_tuned(p, dtype, batch, M, N, G)) return p;
  const int64_t M16 = (M + 15) / 16, N16 = (N + 15) / 16;
  int64_t best = -1;
  for (int bm = 1; bm <= std::min<int64_t>(M16, kBlkRows16 - 1); ++bm)
    for (int bn = 1; bn <= std::min<int64_t>(N16, kBlkRows16 - bm); ++bn) {
      if (bm * bn > blk_tiles_max(dtype)) continue;
      const int64_t mb = (M16 + bm - 1) / bm, nb = (N16 + bn - 1) / bn;
      const int64_t c = tiling_cost(mb * nb * bm * bn, mb * nb * (bm + bn));
      if (best < 0 || c < best) {
        best = c;
        p.bm16 = bm;
        p.bn16 = bn;
        p.mtiles = (int)mb;
        p.ntiles = (int)nb;
      }
    }
  const int64_t mt = (M + kTile - 1) / kTile, nt = (N + kTile - 1) / kTile;
  p.block = best <= tiling_cost(mt * nt * 16, mt * nt * 8);
  if (!p.block) {
    p.mtiles = (int)mt;
    p.ntiles = (int)nt;
  }
  const int64_t tiles = std::max<int64_t>(1, batch * p.mtiles * p.ntiles);
  const int64_t target = p.block ? kBlkTarget : kTileTarget;
  const int64_t want = std::max<int64_t>(1, (target + tiles - 1) / tiles);
  const int64_t most = std::max<int64_t>(1, (G + kMinChunk - 1) / kMinChunk);
  dot_chunks(p, G, std::min<int64_t>(want, most));
  return p;
}

template <typename T, int kMaxT>
int launch_block(const DotPlan& p, dim3 grid, hipStream_t st, const void* A, int64_t lda,
                 int64_t sa, const void* B, int64_t ldb, int64_t sb, int M, int N, int64_t G,
                 void* ws) {
  using C2 = typename Cplx<T>::T;
  static bool lds_set = false;  // the same value from any thread
  if (!lds_set) {
    const hipError_t e = hipFuncSetAttribute(
        reinterpret_cast<const void*>(&k_dot_block<T, kMaxT>),
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)blk_lds_bytes<T>(kBlkRows16));
    if (e != hipSuccess) {
      set_error(std::string("qmx_dot: ") + hipGetErrorString(e));
      return QMX_ERR_HIP;
    }
    lds_set = true;
  }
  hipLaunchKernelGGL((k_dot_block<T, kMaxT>), grid, dim3(kBlkThreads),
                     blk_lds_bytes<T>(p.bm16 + p.bn16), st, static_cast<const C2*>(A), lda, sa,
                     static_cast<const C2*>(B), ldb, sb, M, N, G, p.chunk_len, p.bm16, p.bn16,
                     p.mtiles, p.ntiles, p.nchunks, static_cast<C2*>(ws));
  return QMX_OK;
}

template <typename T>
int dot_impl(const void* A, int64_t lda, int64_t sa, const void* B, int64_t ldb, int64_t sb,
             int64_t batch, int64_t M, int64_t N, int64_t G, void* out, int64_t ldc, int64_t sc,
             int accumulate, void* ws, int64_t ws_bytes, hipStream_t st, bool profile) {
  using C2 = typename Cplx<T>::T;
  const DotPlan p = dot_plan(sizeof(T) == 4 ? QMX_C64 : QMX_C128, batch, M, N, G);
  const int64_t need = batch * p.nchunks * M * N * (int64_t)sizeof(C2);
  QMX_CHECK_ARG(ws && ws_bytes >= need, "qmx_dot: workspace too small");
  QMX_CHECK_ARG(batch * p.mtiles * p.ntiles < (1ll << 31), "qmx_dot: too many tiles");
  hipEvent_t ev = nullptr;
  if (profile) prof_begin(st, &ev);  // events only while profiling is on
  dim3 grid((unsigned)(batch * p.mtiles * p.ntiles), (unsigned)p.nchunks);
  if (p.block) {
    auto go = [&](auto kmax) {
      return launch_block<T, decltype(kmax)::value>(p, grid, st, A, lda, sa, B, ldb, sb, (int)M,
                                                    (int)N, G, ws);
    };
    int rc;
    switch ((p.bm16 * p.bn16 + kBlkThreads / 64 - 1) / (kBlkThreads / 64)) {
      case 1: rc = go(std::integral_constant<int, 1>{}); break;
      case 2: rc = go(std::integral_constant<int, 2>{}); break;
      case 3: rc = go(std::integral_constant<int, 3>{}); break;
      case 4: rc = go(std::integral_constant<int, 4>{}); break;
      case 5: rc = go(std::integral_constant<int, 5>{}); break;
      default: rc = go(std::integral_constant<int, 6>{}); break;
    }
    if (rc) return rc;
  } else {
    hipLaunchKernelGGL((k_dot_partial<T>), grid, dim3(kDotThreads), 0, st,
                       static_cast<const C2*>(A), lda, sa, static_cast<const C2*>(B), ldb, sb,
                       (int)M, (int)N, G, p.chunk_len, p.mtiles, p.ntiles, p.nchunks,
                       static_cast<C2*>(ws));
  }
  QMX_LAUNCHED();
  const int64_t total = batch * M * N;
  hipLaunchKernelGGL((k_dot_reduce<T>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st,
                     static_cast<const C2*>(ws), batch, (int)M, (int)N, p.nchunks,
                     static_cast<double2*>(out), ldc, sc, accumulate);
  QMX_LAUNCHED();
  prof_end_flops(st, ev, QMX_KERNEL_GEMM + (sizeof(T) == 8 ? QMX_KERNEL_C128 : 0),
                 (double)batch * (M + N) * G * sizeof(C2), 8.0 * (double)batch * M * N * G);
  return QMX_OK;
}

}  // namespace

extern "C" {

int64_t qmx_dot_workspace_bytes(int dtype, int64_t batch, int64_t M, int64_t N, int64_t G) {
  if ((dtype != QMX_C64 && dtype != QMX_C128) || batch < 0 || M < 0 || N < 0 || G < 0) return -1;
  const DotPlan p = dot_plan(dtype, batch, M, N, G);
  return std::max<int64_t>(1, batch * p.nchunks * M * N * (dtype == QMX_C64 ? 8 : 16));
}

static int dot_entry(int dtype, const void* A, int64_t lda, int64_t sa, const void* B,
                     int64_t ldb, int64_t sb, int64_t batch, int64_t M, int64_t N, int64_t G,
                     void* out, int64_t ldc, int64_t sc, int accumulate, void* workspace,
                     int64_t workspace_bytes, void* stream, bool profile) {
  QMX_CHECK_ARG(out, "qmx_dot: null output");
  QMX_CHECK_ARG(batch >= 0 && M >= 0 && N >= 0 && G >= 0, "qmx_dot: negative size");
  QMX_CHECK_ARG(M < (1 << 20) && N < (1 << 20), "qmx_dot: M, N too large");
  if (batch == 0 || M == 0 || N == 0) return QMX_OK;
  QMX_CHECK_ARG(A && B, "qmx_dot: null operand");
  QMX_CHECK_ARG(lda >= G && ldb >= G && ldc >= N, "qmx_dot: leading dimension too small");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == QMX_C64)
    return dot_impl<float>(A, lda, sa, B, ldb, sb, batch, M, N, G, out, ldc, sc, accumulate,
                           workspace, workspace_bytes, st, profile);
  if (dtype == QMX_C128)
    return dot_impl<double>(A, lda, sa, B, ldb, sb, batch, M, N, G, out, ldc, sc, accumulate,
                            workspace, workspace_bytes, st, profile);
  set_error("qmx_dot: dtype must be QMX_C64 or QMX_C128");
  return QMX_ERR_ARG;
}

int qmx_dot(int dtype, const void* A, int64_t lda, int64_t sa, const void* B, int64_t ldb,
            int64_t sb, int64_t batch, int64_t M, int64_t N, int64_t G, void* out, int64_t ldc,
            int64_t sc, int accumulate, void* workspace, int64_t workspace_bytes, void* stream) {
  return dot_entry(dtype, A, lda, sa, B, ldb, sb, batch, M, N, G, out, ldc, sc, accumulate,
                   workspace, workspace_bytes, stream, true);
}

// ------------------------------------------------------------ no allocation
// qmx_dot_noalloc keeps its chunk partials in a fixed device array instead
// of a caller workspace. Each GPU holds its own copy of the array for the
// process lifetime; host code never uses its name as an address: the
// device address comes from hipGetSymbolAddress, once per device, with that
// device current.
constexpr int64_t kDotScratchBytes = int64_t(32) << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kDotScratchBytes];
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices];

static int dot_scratch(void** out) {
  int dev = 0;
  QMX_HIP(hipGetDevice(&dev));
  QMX_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "qmx_dot_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    QMX_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    QMX_CHECK_ARG(ptr != nullptr, "qmx_dot_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return QMX_OK;
}

int64_t qmx_dot_noalloc_max_bytes(void) { return kDotScratchBytes; }

int qmx_dot_noalloc(int dtype, const void* A, int64_t lda, int64_t sa, const void* B, int64_t ldb,
                    int64_t sb, int64_t batch, int64_t M, int64_t N, int64_t G, void* out,
                    int64_t ldc, int64_t sc, int accumulate, void* stream) {
  const int64_t need = qmx_dot_workspace_bytes(dtype, batch, M, N, G);
  QMX_CHECK_ARG(need >= 0, "qmx_dot_noalloc: bad dtype or negative size");
  if (need > kDotScratchBytes) {
    set_error("qmx_dot_noalloc: the chunk partials need " + std::to_string(need) +
              " bytes, more than its fixed scratch of " + std::to_string(kDotScratchBytes) +
              " bytes (use qmx_dot with a workspace)");
    return QMX_ERR_NOMEM;
  }
  void* ws = nullptr;
  const int rc = dot_scratch(&ws);
  if (rc) return rc;
  return dot_entry(dtype, A, lda, sa, B, ldb, sb, batch, M, N, G, out, ldc, sc, accumulate, ws,
                   kDotScratchBytes, stream, false);
}

}  // extern "C"
