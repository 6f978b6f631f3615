// Band-space rotations and back-projections on the matrix cores:
//   out[b][i][g] = alpha sum_k op(M)[b][i][k] W[b][k][g] + beta out[b][i][g]
// with op(M) = M or M^T (M small: bands x bands or projectors x bands) and
// W the long operand (bands or projectors along k, plane waves along g,
// unit stride in g). These are the subspace rotations C V of
// Wavefunction.__matmul__ (dft/electrons/_wavefunction_dot.py:138-191), the
// nonlocal back-projection beta (D beta^H C) (_hamiltonian.py:21-24), the
// Davidson rotations (_davidson.py:212-215) and the Gram-Schmidt update
// R -= C (C^H R).
//
// Layout of the work: a workgroup of four waves owns 64 consecutive plane
// waves (16 per wave, the MFMA column count) and a block of 16 MT output
// rows. op(M) is staged through LDS in slabs along k and shared by the
// waves; each wave loads its W operand straight from global memory (four
// rows of 16 consecutive plane waves per load, 128-byte segments) and
// accumulates MT 16 x 16 output tiles. So W is read once per row block:
// once in all when the output rows fit one block (every rotation of the
// batched solves), which makes the product bandwidth-bound where the
// arithmetic type allows it.
//
// Arithmetic:
// - k_rotate<float>: FP32 MFMA 16x16x4, three real products per complex
//   one (Gauss: Mr Wr, Mi Wi and (Mr + Mi)(Wr + Wi)), which adds at most one
//   rounding of the operand sums and doubles the error bound of the
//   imaginary part against the four-product form;
// - k_rotate<double>: FP64 MFMA 16x16x4 (the plain FP64 form), the same
//   three products;
// - k_rotate_bf16x3: complex64 data as BF16 pairs (hi = RNE(x),
//   lo = RNE(x - hi)), three BF16 products per real product
//   (hi hi + hi lo + lo hi) on MFMA 16x16x32 with FP32 accumulation. Each
//   operand keeps ~16 of float32's 24 mantissa bits: relative error about
//   2^-16 per product, 16x the BF16 MFMA rate of FP32 MFMA divided by 3.
#include <algorithm>
#include <string>

#include "qmx_common.h"
#include "qmx_rsrc.h"
#include "qmx_tune.h"

namespace qmx {
void prof_begin(hipStream_t st, hipEvent_t* a);
void prof_end_flops(hipStream_t st, hipEvent_t a, int id, double bytes, double flops);
}  // namespace qmx

using namespace qmx;

namespace {

typedef float rot_f4 __attribute__((ext_vector_type(4)));
typedef double rot_d4 __attribute__((ext_vector_type(4)));
typedef __bf16 rot_bf8 __attribute__((ext_vector_type(8)));

constexpr int kRotThreads = 256;  // four waves
constexpr int kRotCols = 64;      // plane waves per workgroup (16 per wave)
// k per slab. 32 was measured slower for complex64 (62 against 82 TFLOP/s
// on 128 x (144 -> 72) over 3456 plane waves): its prefetched operands take
// the registers that keep more waves per SIMD resident, and those waves
// hide the load latency.
constexpr int kRotSlab = 16;

template <typename T>
struct RotMfma;
// lane l: A[l % 16][l / 16], B[l / 16][l % 16]; result register r of lane l
// holds row row(l / 16, r), column l % 16.
template <>
struct RotMfma<float> {
  using Acc = rot_f4;
  __device__ static int row(int lk, int r) { return 4 * lk + r; }
  __device__ static Acc mac(float a, float b, Acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};
template <>
struct RotMfma<double> {
  using Acc = rot_d4;
  __device__ static int row(int lk, int r) { return lk + 4 * r; }
  __device__ static Acc mac(double a, double b, Acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};

// LDS row stride (elements) of a 16-row MFMA block of MT tiles: the four
// k rows a wave reads in one access fall on distinct bank groups.
template <int MT>
constexpr int rot_ldi() { return (16 * MT) % 32 == 0 ? 16 * MT + 16 : 16 * MT; }

// Element (i, k) of op(M) for a batch; M may be complex64 or complex128
// (MC2).
template <typename MC2>
__device__ __forceinline__ MC2 op_m(const MC2* Mb, int64_t ldm, int transM, int i, int k) {
  return transM ? Mb[(int64_t)k * ldm + i] : Mb[(int64_t)i * ldm + k];
}

// One slab is 16 values of k. W and op(M) are read through buffer
// resources (qmx_rsrc.h) whose base advances one slab per slab: lane
// offsets are computed once. W's resource ends at row K, so rows past K
// read zero; op(M) entries past row m or column K get the out-of-range
// offset. Staging map of op(M): thread tid takes k row kk = tid / 16 of the
// slab and rows i = 16 u + tid % 16 (u < MT) of the block, LDS slots at
// fixed offsets from one base. op(M) goes through two LDS buffers: slab
// s + 1 is written while slab s is read, one barrier per slab. The host
// keeps every resource below 2^31 bytes.
template <typename T, typename MC2, int MT>
__global__ void __launch_bounds__(kRotThreads)
    k_rotate(const MC2* __restrict__ M, int64_t ldm, int64_t sm, int transM,
             const typename Cplx<T>::T* __restrict__ W, int64_t ldw, int64_t sw,
             typename Cplx<T>::T* __restrict__ out, int64_t ldo, int64_t so, int m, int K,
             int64_t G, T alpha, T beta) {
  using C2 = typename Cplx<T>::T;
  using MF = RotMfma<T>;
  constexpr int LDI = rot_ldi<MT>();
  constexpr int KS = kRotSlab;
  constexpr int NQ = KS / 4;  // k rows of four per slab
  static_assert(KS == 16, "the staging map covers 16 k per slab");
  __shared__ T sMr[2][KS * LDI], sMi[2][KS * LDI];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int64_t b = blockIdx.z;
  const int i0 = blockIdx.y * 16 * MT;
  const int64_t g = (int64_t)blockIdx.x * kRotCols + 16 * wave + li;
  // Three real products per complex one (Gauss): p1 = Mr Wr, p2 = Mi Wi,
  // p3 = (Mr + Mi)(Wr + Wi); then Re = p1 - p2, Im = p3 - p1 - p2.
  typename MF::Acc p1[MT], p2[MT], p3[MT];
#pragma unroll
  for (int t = 0; t < MT; ++t) {
    p1[t] = typename MF::Acc{T(0), T(0), T(0), T(0)};
    p2[t] = p1[t];
    p3[t] = p1[t];
  }
  // W: lane (column li, k row lk) takes k = k0 + 4 q + lk; columns past G
  // read column G - 1 and are not stored.
  const C2* wslab = W + b * sw;
  int64_t wleft = (int64_t)K * ldw * (int64_t)sizeof(C2);  // bytes from wslab to row K
  const int64_t gc = g < G ? g : G - 1;  // clamped: every load stays inside W's rows
  int woff[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) woff[q] = (int)(((4 * q + lk) * ldw + gc) * (int64_t)sizeof(C2));
  C2 wn[NQ];
  auto load_w = [&]() {
    const auto r = make_rsrc(wslab, wleft);
#pragma unroll
    for (int q = 0; q < NQ; ++q) wn[q] = rsrc_load<C2>(r, woff[q]);
  };
  // op(M): k row kk of the slab, rows 16 u + il of the block.
  const int kk = tid >> 4, il = tid & 15;
  const int64_t m_kstep = transM ? ldm : 1;        // next k
  const int64_t m_istep = transM ? 1 : ldm;        // next row
  const MC2* mslab = M + b * sm;
  int64_t mleft = ((int64_t)(transM ? K : m) * ldm) * (int64_t)sizeof(MC2);
  int moff[MT];
#pragma unroll
  for (int u = 0; u < MT; ++u) {
    const int i = i0 + 16 * u + il;
    moff[u] = i < m ? (int)((i * m_istep + kk * m_kstep) * (int64_t)sizeof(MC2)) : kOutOfRange;
  }
  T mrn[MT], min_[MT];
  auto load_m = [&](int k0) {
    const auto r = make_rsrc(mslab, mleft);
    const bool kok = k0 + kk < K;
#pragma unroll
    for (int u = 0; u < MT; ++u) {
      const MC2 v = rsrc_load<MC2>(r, kok ? moff[u] : kOutOfRange);
      mrn[u] = (T)v.x;
      min_[u] = (T)v.y;
    }
  };
  auto store_m = [&](int buf) {
#pragma unroll
    for (int u = 0; u < MT; ++u) {
      sMr[buf][kk * LDI + 16 * u + il] = mrn[u];
      sMi[buf][kk * LDI + 16 * u + il] = min_[u];
    }
  };
  load_w();
  load_m(0);
  store_m(0);
  __syncthreads();
  int buf = 0;
  for (int k0 = 0; k0 < K; k0 += KS) {
    C2 w[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) w[q] = wn[q];
    if (k0 + KS < K) {  // next slab's operands, in flight during this slab's MFMAs
      const int64_t wstep = (int64_t)KS * ldw * (int64_t)sizeof(C2);
      wslab += (int64_t)KS * ldw;
      wleft -= wstep;
      mslab += KS * m_kstep;
      mleft -= (int64_t)KS * m_kstep * (int64_t)sizeof(MC2);
      load_w();
      load_m(k0 + KS);
    }
    // Each of the three real products runs over all tiles before the next,
    // so consecutive MFMAs write different accumulators.
    const T* Mr = sMr[buf];
    const T* Mi = sMi[buf];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int kr = (4 * q + lk) * LDI + li;
      T ar[MT], ai[MT];
#pragma unroll
      for (int t = 0; t < MT; ++t) {
        ar[t] = Mr[kr + 16 * t];
        ai[t] = Mi[kr + 16 * t];
      }
      const T ws = w[q].x + w[q].y;
#pragma unroll
      for (int t = 0; t < MT; ++t) p1[t] = MF::mac(ar[t], w[q].x, p1[t]);
#pragma unroll
      for (int t = 0; t < MT; ++t) p2[t] = MF::mac(ai[t], w[q].y, p2[t]);
#pragma unroll
      for (int t = 0; t < MT; ++t) p3[t] = MF::mac(ar[t] + ai[t], ws, p3[t]);
    }
    // The other buffer was last read a slab ago, before the last barrier
    // (after the final slab this stores unused values).
    buf ^= 1;
    store_m(buf);
    __syncthreads();
  }
  if (g >= G) return;
  C2* ob = out + b * so;
#pragma unroll
  for (int t = 0; t < MT; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = i0 + 16 * t + MF::row(lk, r);
      if (i < m) {
        C2* o = ob + (int64_t)i * ldo + g;
        C2 v{alpha * (p1[t][r] - p2[t][r]), alpha * (p3[t][r] - p1[t][r] - p2[t][r])};
        if (beta != T(0)) {
          const C2 p = *o;
          v.x += beta * p.x;
          v.y += beta * p.y;
        }
        *o = v;
      }
    }
}

// The same product with 64-bit addressing throughout, for operands past
// the buffer-resource range (one batch of W or op(M) of 2^30 bytes or more):
// op(M) staged in one LDS buffer (two barriers per slab), W rows clamped to
// K - 1 in the last slab and masked where used.
template <typename T, typename MC2, int MT, int KSL = kRotSlab>
__global__ void __launch_bounds__(kRotThreads)
    k_rotate_flat(const MC2* __restrict__ M, int64_t ldm, int64_t sm, int transM,
             const typename Cplx<T>::T* __restrict__ W, int64_t ldw, int64_t sw,
             typename Cplx<T>::T* __restrict__ out, int64_t ldo, int64_t so, int m, int K,
             int64_t G, T alpha, T beta) {
  using C2 = typename Cplx<T>::T;
  using MF = RotMfma<T>;
  constexpr int LDI = rot_ldi<MT>();
  constexpr int ROWS = 16 * MT;
  constexpr int KS = KSL;  // k per slab
  constexpr int NQ = KS / 4;        // k rows of four per slab
  constexpr int NU = MT * KS / 16;  // op(M) values a thread stages per slab
  static_assert(KS % 16 == 0, "slab of whole 16-row groups");
  __shared__ T sMr[KS * LDI], sMi[KS * LDI];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int64_t b = blockIdx.z;
  const int i0 = blockIdx.y * ROWS;
  const int64_t g = (int64_t)blockIdx.x * kRotCols + 16 * wave + li;
  const MC2* Mb = M + b * sm;
  const C2* Wb = W + b * sw;
  const int64_t gc = g < G ? g : G - 1;  // clamped: loads stay in bounds
  // Three real products per complex one (Gauss): p1 = Mr Wr, p2 = Mi Wi,
  // p3 = (Mr + Mi)(Wr + Wi); then Re = p1 - p2, Im = p3 - p1 - p2.
  typename MF::Acc p1[MT], p2[MT], p3[MT];
#pragma unroll
  for (int t = 0; t < MT; ++t) {
    p1[t] = typename MF::Acc{T(0), T(0), T(0), T(0)};
    p2[t] = p1[t];
    p3[t] = p1[t];
  }
  // W operand of a slab: lane (column li, k row lk) takes k = k0 + 4 q + lk.
  // Loads are clamped in bounds and masked where used. The next slab's W
  // and op(M) values are issued before this slab's barriers, so their
  // latency overlaps this slab's MFMAs; op(M) element u of a thread is
  // (row i, k kk) with tid + kRotThreads u = kk ROWS + i.
  C2 wn[NQ];
  T mrn[NU], min_[NU];
  auto load_w = [&](int k0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int k = k0 + 4 * q + lk;
      wn[q] = Wb[(int64_t)(k < K ? k : K - 1) * ldw + gc];
    }
  };
  auto load_m = [&](int k0) {
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int e = tid + kRotThreads * u;  // < KS * ROWS
      const int kk = e / ROWS, i = e - kk * ROWS;
      const int gi = i0 + i, gk = k0 + kk;
      mrn[u] = T(0);
      min_[u] = T(0);
      if (gi < m && gk < K) {
        const MC2 v = op_m(Mb, ldm, transM, gi, gk);
        mrn[u] = (T)v.x;
        min_[u] = (T)v.y;
      }
    }
  };
  load_w(0);
  load_m(0);
  for (int k0 = 0; k0 < K; k0 += KS) {
    C2 w[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const bool ok = k0 + 4 * q + lk < K;
      w[q] = ok ? wn[q] : C2{T(0), T(0)};
    }
    T mr[NU], mi[NU];
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      mr[u] = mrn[u];
      mi[u] = min_[u];
    }
    if (k0 + KS < K) {
      load_w(k0 + KS);
      load_m(k0 + KS);
    }
    __syncthreads();  // the previous slab's LDS reads are done
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int e = tid + kRotThreads * u;
      const int kk = e / ROWS, i = e - kk * ROWS;
      sMr[kk * LDI + i] = mr[u];
      sMi[kk * LDI + i] = mi[u];
    }
    __syncthreads();
    // Each of the three real products runs over all tiles before the next,
    // so consecutive MFMAs write different accumulators.
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int kr = (4 * q + lk) * LDI + li;
      T ar[MT], ai[MT];
#pragma unroll
      for (int t = 0; t < MT; ++t) {
        ar[t] = sMr[kr + 16 * t];
        ai[t] = sMi[kr + 16 * t];
      }
      const T ws = w[q].x + w[q].y;
#pragma unroll
      for (int t = 0; t < MT; ++t) p1[t] = MF::mac(ar[t], w[q].x, p1[t]);
#pragma unroll
      for (int t = 0; t < MT; ++t) p2[t] = MF::mac(ai[t], w[q].y, p2[t]);
#pragma unroll
      for (int t = 0; t < MT; ++t) p3[t] = MF::mac(ar[t] + ai[t], ws, p3[t]);
    }
  }
  if (g >= G) return;
  C2* ob = out + b * so;
#pragma unroll
  for (int t = 0; t < MT; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = i0 + 16 * t + MF::row(lk, r);
      if (i < m) {
        C2* o = ob + (int64_t)i * ldo + g;
        C2 v{alpha * (p1[t][r] - p2[t][r]), alpha * (p3[t][r] - p1[t][r] - p2[t][r])};
        if (beta != T(0)) {
          const C2 p = *o;
          v.x += beta * p.x;
          v.y += beta * p.y;
        }
        *o = v;
      }
    }
}

// ----------------------------------------------------------- BF16 x 3
// Round to nearest even of a float to its top 16 bits (the BF16 bits).
__device__ __forceinline__ uint32_t bf16_rne(float x) {
  const uint32_t u = __float_as_uint(x);
  return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}
// x = hi + lo with hi, lo BF16 (as 16-bit patterns).
__device__ __forceinline__ void bf16_split(float x, uint32_t& hi, uint32_t& lo) {
  hi = bf16_rne(x);
  lo = bf16_rne(x - __uint_as_float(hi << 16));
}

union Bf8 {
  rot_bf8 v;
  uint32_t u[4];
};

template <typename MC2, int MT>
__global__ void __launch_bounds__(kRotThreads)
    k_rotate_bf16x3(const MC2* __restrict__ M, int64_t ldm, int64_t sm, int transM,
                    const float2* __restrict__ W, int64_t ldw, int64_t sw,
                    float2* __restrict__ out, int64_t ldo, int64_t so, int m, int K, int64_t G,
                    float alpha, float beta) {
  constexpr int ROWS = 16 * MT;
  constexpr int KS = 32;     // k per slab = one MFMA 16x16x32
  constexpr int RS = KS + 8;  // LDS row stride in BF16 elements (80 bytes)
  // Planes: re hi, re lo, im hi, im lo of op(M)[i][k], k contiguous.
  __shared__ __align__(16) uint16_t sM[4][ROWS * RS];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int64_t b = blockIdx.z;
  const int i0 = blockIdx.y * ROWS;
  const int64_t g = (int64_t)blockIdx.x * kRotCols + 16 * wave + li;
  const MC2* Mb = M + b * sm;
  const float2* Wb = W + b * sw;
  const int64_t gc = g < G ? g : G - 1;
  rot_f4 re[MT], im[MT];
#pragma unroll
  for (int t = 0; t < MT; ++t) {
    re[t] = rot_f4{0.f, 0.f, 0.f, 0.f};
    im[t] = re[t];
  }
  // W operand: lane (column li, k group lk) takes k = k0 + 8 lk + e; the
  // next slab's loads go out before this slab's barriers.
  float2 wn[8];
  auto load_w = [&](int k0) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int k = k0 + 8 * lk + e;
      wn[e] = Wb[(int64_t)(k < K ? k : K - 1) * ldw + gc];
    }
  };
  load_w(0);
  for (int k0 = 0; k0 < K; k0 += KS) {
    float2 w[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) w[e] = k0 + 8 * lk + e < K ? wn[e] : float2{0.f, 0.f};
    if (k0 + KS < K) load_w(k0 + KS);
    __syncthreads();
    for (int e = tid; e < ROWS * KS; e += kRotThreads) {
      const int i = e / KS, kk = e - i * KS;
      const int gi = i0 + i, gk = k0 + kk;
      float vr = 0.f, vi = 0.f;
      if (gi < m && gk < K) {
        const MC2 v = op_m(Mb, ldm, transM, gi, gk);
        vr = (float)v.x;
        vi = (float)v.y;
      }
      uint32_t h, l;
      bf16_split(vr, h, l);
      sM[0][i * RS + kk] = (uint16_t)h;
      sM[1][i * RS + kk] = (uint16_t)l;
      bf16_split(vi, h, l);
      sM[2][i * RS + kk] = (uint16_t)h;
      sM[3][i * RS + kk] = (uint16_t)l;
    }
    // Split W while the staging completes: Wr, Wi and -Wi as hi / lo.
    Bf8 wrh, wrl, wih, wil, wnh, wnl;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      uint32_t h0, l0, h1, l1;
      bf16_split(w[2 * e].x, h0, l0);
      bf16_split(w[2 * e + 1].x, h1, l1);
      wrh.u[e] = h0 | (h1 << 16);
      wrl.u[e] = l0 | (l1 << 16);
      bf16_split(w[2 * e].y, h0, l0);
      bf16_split(w[2 * e + 1].y, h1, l1);
      wih.u[e] = h0 | (h1 << 16);
      wil.u[e] = l0 | (l1 << 16);
      wnh.u[e] = wih.u[e] ^ 0x80008000u;
      wnl.u[e] = wil.u[e] ^ 0x80008000u;
    }
    __syncthreads();
    // Tiles in pairs; the twelve BF16 products of a pair alternate between
    // its four accumulators:
    //   re += Mr Wr - Mi Wi,  im += Mr Wi + Mi Wr  (each as hh + hl + lh).
#pragma unroll
    for (int t0 = 0; t0 < MT; t0 += 2) {
      constexpr int kPair = 2;
      Bf8 mrh[kPair], mrl[kPair], mih[kPair], mil[kPair];
#pragma unroll
      for (int p = 0; p < kPair; ++p) {
        const int t = t0 + p < MT ? t0 + p : t0;
        const int off = (16 * t + li) * RS + 8 * lk;
        mrh[p].v = *reinterpret_cast<const rot_bf8*>(&sM[0][off]);
        mrl[p].v = *reinterpret_cast<const rot_bf8*>(&sM[1][off]);
        mih[p].v = *reinterpret_cast<const rot_bf8*>(&sM[2][off]);
        mil[p].v = *reinterpret_cast<const rot_bf8*>(&sM[3][off]);
      }
      const int np = t0 + 1 < MT ? 2 : 1;
#define QMX_BF(A, B, ACC, p) ACC[t0 + p] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[p].v, B.v, ACC[t0 + p], 0, 0, 0)
#define QMX_BF2(A, B, ACC) QMX_BF(A, B, ACC, 0); if (np > 1) { QMX_BF(A, B, ACC, 1); }
      QMX_BF2(mrh, wrh, re);
      QMX_BF2(mrh, wih, im);
      QMX_BF2(mrh, wrl, re);
      QMX_BF2(mrh, wil, im);
      QMX_BF2(mrl, wrh, re);
      QMX_BF2(mrl, wih, im);
      QMX_BF2(mih, wnh, re);
      QMX_BF2(mih, wrh, im);
      QMX_BF2(mih, wnl, re);
      QMX_BF2(mih, wrl, im);
      QMX_BF2(mil, wnh, re);
      QMX_BF2(mil, wrh, im);
#undef QMX_BF2
#undef QMX_BF
    }
  }
  if (g >= G) return;
  float2* ob = out + b * so;
#pragma unroll
  for (int t = 0; t < MT; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = i0 + 16 * t + 4 * lk + r;
      if (i < m) {
        float2* o = ob + (int64_t)i * ldo + g;
        float2 v{alpha * re[t][r], alpha * im[t][r]};
        if (beta != 0.f) {
          const float2 p = *o;
          v.x += beta * p.x;
          v.y += beta * p.y;
        }
        *o = v;
      }
    }
}

enum Variant { kFp = 0, kBf16x3 = 1 };

// Output rows per block: the fewest blocks, then the smallest MT that holds
// the rows of one block (W is read once per block).
template <int MTMAX>
int pick_mt(int m, int* blocks) {
  const int nb = (m + 16 * MTMAX - 1) / (16 * MTMAX);
  const int rows = (m + nb - 1) / nb;
  *blocks = nb;
  return std::max(1, (rows + 15) / 16);
}

template <typename T, typename MC2, int MT>
void launch_fp(dim3 grid, hipStream_t st, const void* M, int64_t ldm, int64_t sm, int transM,
               const void* W, int64_t ldw, int64_t sw, void* out, int64_t ldo, int64_t so, int m,
               int K, int64_t G, double alpha, double beta) {
  using C2 = typename Cplx<T>::T;
  // Buffer resources while one batch of W and of op(M), with the lanes'
  // offsets past them, stays below 2^30 bytes.
  const bool rsrc = (int64_t)std::max(K, 16) * ldw * (int64_t)sizeof(C2) < (1ll << 30) &&
                    (int64_t)(K + m + 256) * ldm * (int64_t)sizeof(MC2) < (1ll << 30);
  if (rsrc)
    hipLaunchKernelGGL((k_rotate<T, MC2, MT>), grid, dim3(kRotThreads), 0, st,
                       static_cast<const MC2*>(M), ldm, sm, transM, static_cast<const C2*>(W),
                       ldw, sw, static_cast<C2*>(out), ldo, so, m, K, G, (T)alpha, (T)beta);
  else
    hipLaunchKernelGGL((k_rotate_flat<T, MC2, MT>), grid, dim3(kRotThreads), 0, st,
                       static_cast<const MC2*>(M), ldm, sm, transM, static_cast<const C2*>(W),
                       ldw, sw, static_cast<C2*>(out), ldo, so, m, K, G, (T)alpha, (T)beta);
}

template <typename MC2, int MT>
void launch_bf(dim3 grid, hipStream_t st, const void* M, int64_t ldm, int64_t sm, int transM,
               const void* W, int64_t ldw, int64_t sw, void* out, int64_t ldo, int64_t so, int m,
               int K, int64_t G, double alpha, double beta) {
  hipLaunchKernelGGL((k_rotate_bf16x3<MC2, MT>), grid, dim3(kRotThreads), 0, st,
                     static_cast<const MC2*>(M), ldm, sm, transM, static_cast<const float2*>(W),
                     ldw, sw, static_cast<float2*>(out), ldo, so, m, K, G, (float)alpha,
                     (float)beta);
}

// In place (out == W): a workgroup reads every row of W at its columns
// before it writes any output row there, and no other workgroup touches
// those columns, provided the output rows form one block (grid.y == 1).
template <typename T, typename MC2>
int rotate_dispatch(int variant, int transM, int64_t batch, int m, int K, int64_t G,
                    double alpha, const void* M, int64_t ldm, int64_t sm, const void* W,
                    int64_t ldw, int64_t sw, double beta, void* out, int64_t ldo, int64_t so,
                    hipStream_t st) {
  // FP64 accumulators take twice the registers: at most 5 tiles per wave.
  constexpr int kMaxFp = sizeof(T) == 8 ? 5 : 8;
  const bool inplace = out == W;
  int blocks = 1;
  int mt = variant == kBf16x3 ? pick_mt<6>(m, &blocks) : pick_mt<kMaxFp>(m, &blocks);
  if (inplace) {
    QMX_CHECK_ARG(ldo == ldw && so == sw, "qmx_gemm: in place needs out and W with one layout");
    QMX_CHECK_ARG(blocks == 1, "qmx_gemm: in place needs m <= 16 x the tile limit");
  } else {
    const int64_t key[6] = {sizeof(T) == 4 ? QMX_C64 : QMX_C128,
                            sizeof(MC2) == 8 ? QMX_C64 : QMX_C128, variant, m, K, tune_g_bucket(G)};
    int64_t ch[1];
    const int mt_max = variant == kBf16x3 ? 6 : kMaxFp;
    if (tune_lookup(QMX_TUNE_ROTATE, key, 6, ch, 1) && ch[0] >= 1 && ch[0] <= mt_max) {
      mt = (int)ch[0];
      blocks = (m + 16 * mt - 1) / (16 * mt);
    }
  }
  const int64_t gtiles = (G + kRotCols - 1) / kRotCols;
  QMX_CHECK_ARG(gtiles < (1ll << 31) && blocks < 65536 && batch < 65536,
                "qmx_gemm: grid too large");
  dim3 grid((unsigned)gtiles, (unsigned)blocks, (unsigned)batch);
  hipEvent_t ev = nullptr;
  prof_begin(st, &ev);
  if (variant == kBf16x3) {
    if constexpr (sizeof(T) == 4) {
      switch (mt) {
        case 1: launch_bf<MC2, 1>(grid, st, M, ldm, sm, transM, W, ldw, sw, out, ldo, so, m, K, G, alpha, beta); break;
        case 2: launch_bf<MC2, 2>(grid, st, M, ldm, sm, transM, W, ldw, sw, out, ldo, so, m, K, G, alpha, beta); break;
        case 3: launch_bf<MC2, 3>(grid, st, M, ldm, sm, transM, W, ldw, sw, out, ldo, so, m, K, G, alpha, beta); break;
        case 4: launch_bf<MC2, 4>(grid, st, M, ldm, sm, transM, W, ldw, sw, out, ldo, so, m, K, G, alpha, beta); break;
        case 5: launch_bf<MC2, 5>(grid, st, M, ldm, sm, transM, W, ldw, sw, out, ldo, so, m, K, G, alpha, beta); break;
        default: launch_bf<MC2, 6>(grid, st, M, ldm, sm, transM, W, ldw, sw, out, ldo, so, m, K, G, alpha, beta); break;
      }
    }
  } else {
    switch (mt) {
      case 1: launch_fp<T, MC2, 1>(grid, st, M, ldm, sm, transM, W, ldw, sw, out, ldo, so, m, K, G, alpha, beta); break;
      case 2: launch_fp<T, MC2, 2>(grid, st, M, ldm, sm, transM, W, ldw, sw, out, ldo, so, m, K, G, alpha, beta); break;
      case 3: launch_fp<T, MC2, 3>(grid, st, M, ldm, sm, transM, W, ldw, sw, out, ldo, so, m, K, G, alpha, beta); break;
      case 4: launch_fp<T, MC2, 4>(grid, st, M, ldm, sm, transM, W, ldw, sw, out, ldo, so, m, K, G, alpha, beta); break;
      case 5: launch_fp<T, MC2, 5>(grid, st, M, ldm, sm, transM, W, ldw, sw, out, ldo, so, m, K, G, alpha, beta); break;
      case 6: if constexpr (kMaxFp >= 6) { launch_fp<T, MC2, 6>(grid, st, M, ldm, sm, transM, W, ldw, sw, out, ldo, so, m, K, G, alpha, beta); } break;
      case 7: if constexpr (kMaxFp >= 7) { launch_fp<T, MC2, 7>(grid, st, M, ldm, sm, transM, W, ldw, sw, out, ldo, so, m, K, G, alpha, beta); } break;
      default: if constexpr (kMaxFp >= 8) { launch_fp<T, MC2, 8>(grid, st, M, ldm, sm, transM, W, ldw, sw, out, ldo, so, m, K, G, alpha, beta); } break;
    }
  }
  QMX_LAUNCHED();
  // Compulsory bytes: W read once per row block, out written (and read when
  // beta != 0), the small matrix once per batch.
  const double cs = (double)sizeof(typename Cplx<T>::T);
  const double by = (double)batch * G * cs * ((double)K * blocks + m * (beta != 0.0 ? 2.0 : 1.0)) +
                    (double)batch * m * K * sizeof(MC2);
  const int id = (variant == kBf16x3 ? QMX_KERNEL_ROTATE_BF16X3 : QMX_KERNEL_ROTATE) +
                 (sizeof(T) == 8 ? QMX_KERNEL_C128 : 0);
  // Algorithmic flops: one complex multiply-add (8 real flops) per (b, i, k, g).
  prof_end_flops(st, ev, id, by, 8.0 * (double)batch * m * K * G);
  return QMX_OK;
}

}  // namespace

extern "C" {

int qmx_gemm(int dtype, int m_dtype, int variant, int trans_m, int64_t batch, int64_t m,
             int64_t K, int64_t G, double alpha, const void* M, int64_t ldm, int64_t sm,
             const void* W, int64_t ldw, int64_t sw, double beta, void* out, int64_t ldo,
             int64_t so, void* stream) {
  QMX_CHECK_ARG(batch >= 0 && m >= 0 && K >= 0 && G >= 0, "qmx_gemm: negative size");
  QMX_CHECK_ARG(m < (1 << 20) && K < (1 << 20), "qmx_gemm: m or K too large");
  QMX_CHECK_ARG(dtype == QMX_C64 || dtype == QMX_C128, "qmx_gemm: dtype must be QMX_C64 or QMX_C128");
  QMX_CHECK_ARG(m_dtype == QMX_C64 || m_dtype == QMX_C128, "qmx_gemm: bad matrix dtype");
  QMX_CHECK_ARG(variant == kFp || variant == kBf16x3, "qmx_gemm: variant must be 0 or 1");
  QMX_CHECK_ARG(variant == kFp || dtype == QMX_C64, "qmx_gemm: BF16x3 needs complex64 data");
  if (batch == 0 || m == 0 || G == 0) return QMX_OK;
  QMX_CHECK_ARG(out, "qmx_gemm: null output");
  QMX_CHECK_ARG(ldo >= G, "qmx_gemm: ldo < G");
  if (K == 0) {
    QMX_CHECK_ARG(beta == 1.0, "qmx_gemm: K = 0 needs beta = 1");
    return QMX_OK;
  }
  QMX_CHECK_ARG(M && W, "qmx_gemm: null operand");
  QMX_CHECK_ARG(ldw >= G, "qmx_gemm: ldw < G");
  QMX_CHECK_ARG(ldm >= (trans_m ? m : K), "qmx_gemm: ldm too small");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int tm = trans_m ? 1 : 0;
  if (dtype == QMX_C64)
    return m_dtype == QMX_C64
               ? rotate_dispatch<float, float2>(variant, tm, batch, (int)m, (int)K, G, alpha, M,
                                                ldm, sm, W, ldw, sw, beta, out, ldo, so, st)
               : rotate_dispatch<float, double2>(variant, tm, batch, (int)m, (int)K, G, alpha, M,
                                                 ldm, sm, W, ldw, sw, beta, out, ldo, so, st);
  return m_dtype == QMX_C64
             ? rotate_dispatch<double, float2>(variant, tm, batch, (int)m, (int)K, G, alpha, M,
                                               ldm, sm, W, ldw, sw, beta, out, ldo, so, st)
             : rotate_dispatch<double, double2>(variant, tm, batch, (int)m, (int)K, G, alpha, M,
                                                ldm, sm, W, ldw, sw, beta, out, ldo, so, st);
}

// BLAS-shaped entry points (row-major): C (m x n) = alpha op(A) (m x k)
// B (k x n) + beta C, batched with strides; op(A) = A^T when transA, B not
// transposed (its rows are the long axis).
static int gemm_blas(int dtype, int variant, int transA, int transB, int64_t m, int64_t n,
                     int64_t k, double alpha, const void* A, int64_t lda, int64_t sA,
                     const void* B, int64_t ldb, int64_t sB, double beta, void* C, int64_t ldc,
                     int64_t sC, int64_t batch, void* stream) {
  QMX_CHECK_ARG(!transB, "qmx_gemm_*: transposed B is not supported (its rows are the long axis)");
  return qmx_gemm(dtype, dtype, variant, transA, batch, m, k, n, alpha, A, lda, sA, B, ldb, sB,
                  beta, C, ldc, sC, stream);
}

int qmx_gemm_c64(int transA, int transB, int64_t m, int64_t n, int64_t k, double alpha,
                 const void* A, int64_t lda, int64_t sA, const void* B, int64_t ldb, int64_t sB,
                 double beta, void* C, int64_t ldc, int64_t sC, int64_t batch, void* stream) {
  return gemm_blas(QMX_C64, kFp, transA, transB, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta,
                   C, ldc, sC, batch, stream);
}

int qmx_gemm_c64_bf16x3(int transA, int transB, int64_t m, int64_t n, int64_t k, double alpha,
                        const void* A, int64_t lda, int64_t sA, const void* B, int64_t ldb,
                        int64_t sB, double beta, void* C, int64_t ldc, int64_t sC, int64_t batch,
                        void* stream) {
  return gemm_blas(QMX_C64, kBf16x3, transA, transB, m, n, k, alpha, A, lda, sA, B, ldb, sB,
                   beta, C, ldc, sC, batch, stream);
}

int qmx_gemm_c128(int transA, int transB, int64_t m, int64_t n, int64_t k, double alpha,
                  const void* A, int64_t lda, int64_t sA, const void* B, int64_t ldb, int64_t sB,
                  double beta, void* C, int64_t ldc, int64_t sC, int64_t batch, void* stream) {
  return gemm_blas(QMX_C128, kFp, transA, transB, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta,
                   C, ldc, sC, batch, stream);
}

}  // extern "C"
