// LDS-resident batched Stockham FFT core (mixed radix 2, 3, 4, 5).
//
// A workgroup transforms M lines of length n held in LDS with element
// (position p, line l) at buf[p * ld + l]; consecutive threads take
// consecutive lines, so every butterfly pass is a conflict-free sweep of
// LDS. Pass with radix R and sub-transform length Ns (product of the radices
// already applied) maps butterfly j of a line to
//   in:  j + r n/R,       twiddle  w^(r (j mod Ns) n/(Ns R)),  w = e^(-2 pi i/n)
//   out: (j/Ns) Ns R + (j mod Ns) + r Ns
// which leaves the result in natural order after the last pass (the
// autosort Stockham formulation of Govindaraju et al., SC'08). Twiddles
// come from an exact table tw[t] = e^(-2 pi i t/n) staged in LDS.
#pragma once

#include "qmx_common.h"

namespace qmx {

// DFT of size R in registers; S = -1 forward, +1 inverse.
template <typename C2, int S>
__device__ __forceinline__ void dft2(C2& a0, C2& a1) {
  C2 t = a0;
  a0 = cadd(t, a1);
  a1 = csub(t, a1);
}

template <typename C2, int S>
__device__ __forceinline__ void dft4(C2& a0, C2& a1, C2& a2, C2& a3) {
  C2 t0 = cadd(a0, a2), t1 = csub(a0, a2);
  C2 t2 = cadd(a1, a3), t3 = mul_si<C2, S>(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}

template <typename C2, int S>
__device__ __forceinline__ void dft3(C2& a0, C2& a1, C2& a2) {
  using R = decltype(a0.x);
  const R h = R(0.86602540378443864676);  // sqrt(3)/2
  C2 t = cadd(a1, a2);
  C2 m = C2{a0.x - R(0.5) * t.x, a0.y - R(0.5) * t.y};
  C2 d = mul_si<C2, S>(cscale(csub(a1, a2), h));
  a0 = cadd(a0, t);
  a1 = cadd(m, d);
  a2 = csub(m, d);
}

template <typename C2, int S>
__device__ __forceinline__ void dft5(C2& a0, C2& a1, C2& a2, C2& a3, C2& a4) {
  using R = decltype(a0.x);
  const R c1 = R(0.30901699437494742410);   // cos(2 pi/5)
  const R c2 = R(-0.80901699437494742410);  // cos(4 pi/5)
  const R s1 = R(0.95105651629515357212);   // sin(2 pi/5)
  const R s2 = R(0.58778525229247312917);   // sin(4 pi/5)
  C2 b1 = cadd(a1, a4), b2 = cadd(a2, a3);
  C2 d1 = csub(a1, a4), d2 = csub(a2, a3);
  C2 m1 = C2{a0.x + c1 * b1.x + c2 * b2.x, a0.y + c1 * b1.y + c2 * b2.y};
  C2 m2 = C2{a0.x + c2 * b1.x + c1 * b2.x, a0.y + c2 * b1.y + c1 * b2.y};
  C2 e1 = mul_si<C2, S>(C2{s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y});
  C2 e2 = mul_si<C2, S>(C2{s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y});
  a0 = cadd(a0, cadd(b1, b2));
  a1 = cadd(m1, e1);
  a4 = csub(m1, e1);
  a2 = cadd(m2, e2);
  a3 = csub(m2, e2);
}

template <typename C2, int R, int S>
__device__ __forceinline__ void dft(C2 (&v)[R]) {
  if constexpr (R == 2) dft2<C2, S>(v[0], v[1]);
  if constexpr (R == 3) dft3<C2, S>(v[0], v[1], v[2]);
  if constexpr (R == 4) dft4<C2, S>(v[0], v[1], v[2], v[3]);
  if constexpr (R == 5) dft5<C2, S>(v[0], v[1], v[2], v[3], v[4]);
}

// One Stockham pass over the M lines (division by M via fastdiv `fm`).
template <typename C2, int R, int S>
__device__ __forceinline__ void stockham_pass(const C2* __restrict__ a, C2* __restrict__ b,
                                              const C2* __restrict__ tw, int n,
                                              const FastDiv& fns, const FastDiv& fm,
                                              int ld) {
  const int Ns = fns.d;
  const int stride = n / R;
  const int twstep = n / (Ns * R);
  const int total = stride * fm.d;
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    const int j = fdiv(t, fm);
    const int l = t - j * fm.d;
    const int k = j - fdiv(j, fns) * Ns;
    C2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = a[(j + r * stride) * ld + l];
#pragma unroll
    for (int r = 1; r < R; ++r) {
      C2 w = tw[r * k * twstep];
      if constexpr (S > 0) w.y = -w.y;
      v[r] = cmul(v[r], w);
    }
    dft<C2, R, S>(v);
    const int idxD = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) b[(idxD + r * Ns) * ld + l] = v[r];
  }
}

// Transform the M lines in `a` (scratch `b`); returns the buffer holding the
// result. Must be called by the whole workgroup; ends with a barrier.
template <typename C2, int S>
__device__ C2* fft_lds(C2* a, C2* b, const C2* tw, const Radices& rad, const FastDiv& fm,
                       int ld) {
  for (int ip = 0; ip < rad.nr; ++ip) {
    switch (rad.r[ip]) {
      case 4: stockham_pass<C2, 4, S>(a, b, tw, rad.n, rad.ns[ip], fm, ld); break;
      case 2: stockham_pass<C2, 2, S>(a, b, tw, rad.n, rad.ns[ip], fm, ld); break;
      case 3: stockham_pass<C2, 3, S>(a, b, tw, rad.n, rad.ns[ip], fm, ld); break;
      case 5: stockham_pass<C2, 5, S>(a, b, tw, rad.n, rad.ns[ip], fm, ld); break;
    }
    __syncthreads();
    C2* t = a;
    a = b;
    b = t;
  }
  return a;
}

// LDS exchange between the lanes of one wave: all LDS stores of this wave
// are complete before any later LDS load (no workgroup barrier).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// One Stockham pass over the mw lines [l0, l0 + mw) owned by the calling wave.
template <typename C2, int R, int S>
__device__ __forceinline__ void stockham_pass_wave(const C2* __restrict__ a, C2* __restrict__ b,
                                                   const C2* __restrict__ tw, int n,
                                                   const FastDiv& fns, const FastDiv& fmw,
                                                   int ld, int l0, int lane) {
  const int Ns = fns.d;
  const int stride = n / R;
  const int twstep = n / (Ns * R);
  const int total = stride * fmw.d;
  for (int t = lane; t < total; t += 64) {
    const int j = fdiv(t, fmw);
    const int l = l0 + t - j * fmw.d;
    const int k = j - fdiv(j, fns) * Ns;
    C2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = a[(j + r * stride) * ld + l];
#pragma unroll
    for (int r = 1; r < R; ++r) {
      C2 w = tw[r * k * twstep];
      if constexpr (S > 0) w.y = -w.y;
      v[r] = cmul(v[r], w);
    }
    dft<C2, R, S>(v);
    const int idxD = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) b[(idxD + r * Ns) * ld + l] = v[r];
  }
}

// Transform the lines [l0, l0 + mw) of `a` (scratch `b`) with the calling wave
// only; returns the buffer holding the result (the same for every wave).
template <typename C2, int S>
__device__ C2* fft_lds_wave(C2* a, C2* b, const C2* tw, const Radices& rad, const FastDiv& fmw,
                            int ld, int l0) {
  const int lane = threadIdx.x & 63;
  for (int ip = 0; ip < rad.nr; ++ip) {
    switch (rad.r[ip]) {
      case 4: stockham_pass_wave<C2, 4, S>(a, b, tw, rad.n, rad.ns[ip], fmw, ld, l0, lane); break;
      case 2: stockham_pass_wave<C2, 2, S>(a, b, tw, rad.n, rad.ns[ip], fmw, ld, l0, lane); break;
      case 3: stockham_pass_wave<C2, 3, S>(a, b, tw, rad.n, rad.ns[ip], fmw, ld, l0, lane); break;
      case 5: stockham_pass_wave<C2, 5, S>(a, b, tw, rad.n, rad.ns[ip], fmw, ld, l0, lane); break;
    }
    wave_lds_sync();
    C2* t = a;
    a = b;
    b = t;
  }
  return a;
}

// In-place Stockham pass of radix R over the nl lines of a 2D LDS array,
// element (position p, line l) at buf[p * ps + off(l)], off(l) = l * ls, or
// lines[l] * ls when INDIRECT. All NT threads of the block take part; each
// holds at most NB butterflies in registers between the read half and the
// write half, which a barrier separates, so the pass overwrites its own
// input and one buffer suffices. LINE_FAST: consecutive threads take
// consecutive lines (unit stride when ls == 1); otherwise consecutive
// butterflies of one line (unit stride when ps == 1). Callers guarantee
// stride * nl <= NB * NT.
template <typename C2, int R, int S, int NT, int NB, bool LINE_FAST, bool INDIRECT>
__device__ __forceinline__ void ip_pass(C2* buf, const C2* __restrict__ tw, int n, FastDiv fns,
                                        FastDiv fst, int nl, FastDiv fnl, int ps, int ls,
                                        const int* lines) {
  const int Ns = fns.d, stride = fst.d, twstep = n / (Ns * R);
  const int total = stride * nl;
  C2 v[NB][R];
#pragma unroll
  for (int it = 0; it < NB; ++it) {
    const int t = threadIdx.x + it * NT;
    if (t < total) {
      int j, l;
      if (LINE_FAST) {
        j = fdiv(t, fnl);
        l = t - j * nl;
      } else {
        l = fdiv(t, fst);
        j = t - l * stride;
      }
      const int lo = (INDIRECT ? lines[l] : l) * ls;
      const int k = j - fdiv(j, fns) * Ns;
#pragma unroll
      for (int r = 0; r < R; ++r) v[it][r] = buf[(j + r * stride) * ps + lo];
#pragma unroll
      for (int r = 1; r < R; ++r) {
        C2 w = tw[r * k * twstep];
        if constexpr (S > 0) w.y = -w.y;
        v[it][r] = cmul(v[it][r], w);
      }
      dft<C2, R, S>(v[it]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < NB; ++it) {
    const int t = threadIdx.x + it * NT;
    if (t < total) {
      int j, l;
      if (LINE_FAST) {
        j = fdiv(t, fnl);
        l = t - j * nl;
      } else {
        l = fdiv(t, fst);
        j = t - l * stride;
      }
      const int lo = (INDIRECT ? lines[l] : l) * ls;
      const int k = j - fdiv(j, fns) * Ns;
      const int idxD = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[(idxD + r * Ns) * ps + lo] = v[it][r];
    }
  }
  __syncthreads();
}

// Full in-place transform of the nl lines described as for ip_pass; at
// most MAXE complex values per thread in any pass. Ends with a barrier.
template <typename C2, int S, int NT, int MAXE, bool LINE_FAST, bool INDIRECT>
__device__ __forceinline__ void fft_ip(C2* buf, const C2* tw, const Radices& rad, int nl, FastDiv fnl, int ps,
                       int ls, const int* lines) {
  for (int ip = 0; ip < rad.nr; ++ip) {
    switch (rad.r[ip]) {
      case 4:
        ip_pass<C2, 4, S, NT, MAXE / 4, LINE_FAST, INDIRECT>(buf, tw, rad.n, rad.ns[ip],
                                                              rad.st[ip], nl, fnl, ps, ls, lines);
        break;
      case 2:
        ip_pass<C2, 2, S, NT, MAXE / 2, LINE_FAST, INDIRECT>(buf, tw, rad.n, rad.ns[ip],
                                                              rad.st[ip], nl, fnl, ps, ls, lines);
        break;
      case 3:
        ip_pass<C2, 3, S, NT, MAXE / 3, LINE_FAST, INDIRECT>(buf, tw, rad.n, rad.ns[ip],
                                                              rad.st[ip], nl, fnl, ps, ls, lines);
        break;
      case 5:
        ip_pass<C2, 5, S, NT, MAXE / 5, LINE_FAST, INDIRECT>(buf, tw, rad.n, rad.ns[ip],
                                                              rad.st[ip], nl, fnl, ps, ls, lines);
        break;
    }
  }
}

// Largest butterfly count per thread that fft_ip needs for nl lines of the
// factorised length, with NT threads (host side, to check against MAXE / R).
inline bool fft_ip_fits(const Radices& rad, int nl, int nt, int maxe) {
  for (int ip = 0; ip < rad.nr; ++ip) {
    const int R = rad.r[ip];
    const long total = (long)(rad.n / R) * nl;
    if ((total + nt - 1) / nt > maxe / R) return false;
  }
  return true;
}

// Stage the twiddle table of length n into LDS.
template <typename C2>
__device__ __forceinline__ void load_twiddles(C2* dst, const C2* __restrict__ src, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

template <typename C2>
__device__ __forceinline__ void zero_lds(C2* dst, int count) {
  for (int i = threadIdx.x; i < count; i += blockDim.x) dst[i] = C2{0, 0};
}

}  // namespace qmx
