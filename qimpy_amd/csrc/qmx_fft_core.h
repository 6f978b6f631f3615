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

// Composite radix R = R1 R2 (R1, R2 in 2..5) in registers: with n = R2 n1 + n2
// and k = k1 + R1 k2,
//   X[k] = sum_n2 W_R^(n2 k1) W_R2^(n2 k2) sum_n1 v[R2 n1 + n2] W_R1^(n1 k1),
// R2 DFTs of size R1, the internal twiddles W_R^m = tw[m twR] (tw the exact
// n-point table, twR = n / R; conjugated for the inverse), R1 DFTs of size R2.
template <typename C2, int R1, int R2, int S>
__device__ __forceinline__ void dft_comp(C2 (&v)[R1 * R2], const C2* __restrict__ tw, int twR) {
  constexpr int R = R1 * R2;
  C2 y[R];
#pragma unroll
  for (int n2 = 0; n2 < R2; ++n2) {
    C2 a[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) a[n1] = v[R2 * n1 + n2];
    dft<C2, R1, S>(a);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      const int m = (n2 * k1) % R;
      if (m != 0) {
        C2 w = tw[m * twR];
        if constexpr (S > 0) w.y = -w.y;
        a[k1] = cmul(a[k1], w);
      }
      y[n2 * R1 + k1] = a[k1];
    }
  }
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) {
    C2 b[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) b[n2] = y[n2 * R1 + k1];
    dft<C2, R2, S>(b);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) v[k1 + R1 * k2] = b[k2];
  }
}

// Coprime R1, R2 (in 2..5): the prime-factor (Good-Thomas) form, no
// twiddles. Input n = (R2 n1 + R1 n2) mod R, output
// k = (R2 (R2^-1 mod R1) k1 + R1 (R1^-1 mod R2) k2) mod R, so that
// w_R^(n k) = w_R1^(n1 k1) w_R2^(n2 k2): R2 DFTs of size R1, then R1 of
// size R2, all index maps compile-time constants.
constexpr int inv_mod(int a, int m) {
  for (int x = 1; x < m; ++x)
    if ((a * x) % m == 1) return x;
  return 0;
}
constexpr int gcd_int(int a, int b) { return b == 0 ? a : gcd_int(b, a % b); }

template <typename C2, int R1, int R2, int S>
__device__ __forceinline__ void dft_pfa(C2 (&v)[R1 * R2]) {
  constexpr int R = R1 * R2;
  static_assert(gcd_int(R1, R2) == 1, "prime-factor DFT needs coprime factors");
  constexpr int E1 = (R2 * inv_mod(R2, R1)) % R, E2 = (R1 * inv_mod(R1, R2)) % R;
  C2 y[R];
#pragma unroll
  for (int n2 = 0; n2 < R2; ++n2) {
    C2 a[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) a[n1] = v[(R2 * n1 + R1 * n2) % R];
    dft<C2, R1, S>(a);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) y[k1 * R2 + n2] = a[k1];
  }
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) {
    C2 b[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) b[n2] = y[k1 * R2 + n2];
    dft<C2, R2, S>(b);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) v[(E1 * k1 + E2 * k2) % R] = b[k2];
  }
}

// Radices 2..5 directly; composite radices 6, 10, 12, 15 (coprime factors)
// in prime-factor form; 8, 9, 16 through dft_comp with the twiddle table.
template <typename C2, int R, int S>
__device__ __forceinline__ void dft_any(C2 (&v)[R], const C2* __restrict__ tw, int n) {
  if constexpr (R <= 5) {
    dft<C2, R, S>(v);
  } else if constexpr (R == 6) {
    dft_pfa<C2, 2, 3, S>(v);
  } else if constexpr (R == 8) {
    dft_comp<C2, 4, 2, S>(v, tw, n / R);
  } else if constexpr (R == 9) {
    dft_comp<C2, 3, 3, S>(v, tw, n / R);
  } else if constexpr (R == 10) {
    dft_pfa<C2, 2, 5, S>(v);
  } else if constexpr (R == 12) {
    dft_pfa<C2, 4, 3, S>(v);
  } else if constexpr (R == 15) {
    dft_pfa<C2, 3, 5, S>(v);
  } else if constexpr (R == 16) {
    dft_comp<C2, 4, 4, S>(v, tw, n / R);
  }
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
      dft_any<C2, R, S>(v[it], tw, n);
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
#define QMX_IP_CASE(RR)                                                                   \
  case RR:                                                                                \
    if constexpr (MAXE >= RR)                                                             \
      ip_pass<C2, RR, S, NT, MAXE / RR, LINE_FAST, INDIRECT>(buf, tw, rad.n, rad.ns[ip],  \
                                                             rad.st[ip], nl, fnl, ps, ls, \
                                                             lines);                      \
    break;
      QMX_IP_CASE(6)
      QMX_IP_CASE(8)
      QMX_IP_CASE(9)
      QMX_IP_CASE(10)
      QMX_IP_CASE(12)
      QMX_IP_CASE(15)
      QMX_IP_CASE(16)
#undef QMX_IP_CASE
    }
  }
}

// Radices of a length for the in-place passes of nl lines with NT threads
// and MAXE values per thread: the prime-power factors of `in` grouped into
// composite radices of at most 16 (6, 8, 9, 10, 12, 15, 16), the grouping with
// the fewest passes whose butterflies fit the register budget. Falls back
// to `in` when no grouping fits better.
inline Radices combine_radices(const Radices& in, int nl, int nt, int maxe) {
  auto fits = [&](int R) {
    if (R > maxe) return false;
    const long total = (long)(in.n / R) * nl;
    return (total + nt - 1) / nt <= maxe / R;
  };
  auto allowed = [](int R) {
    return R == 2 || R == 3 || R == 4 || R == 5 || R == 6 || R == 8 || R == 9 || R == 10 ||
           R == 12 || R == 15 || R == 16;
  };
  // Multiset of prime factors (a 4 counts as two 2s).
  int f[32], nf = 0;
  for (int i = 0; i < in.nr; ++i) {
    if (in.r[i] == 4) {
      f[nf++] = 2;
      f[nf++] = 2;
    } else {
      f[nf++] = in.r[i];
    }
  }
  // Greedy: repeatedly take the largest allowed product of remaining factors
  // that fits (products of subsets of at most four factors).
  int used[32] = {0};
  int groups[32], ng = 0;
  int left = nf;
  while (left > 0) {
    int best = 0, bestmask = 0;
    const int idx_n = nf;
    for (int mask = 1; mask < (1 << idx_n); ++mask) {
      int prod = 1, cnt = 0;
      bool ok = true;
      for (int i = 0; i < idx_n && ok; ++i) {
        if (!(mask >> i & 1)) continue;
        if (used[i]) ok = false;
        prod *= f[i];
        ++cnt;
        if (prod > 16 || cnt > 4) ok = false;
      }
      if (!ok || !allowed(prod) || !fits(prod)) continue;
      if (prod > best) {
        best = prod;
        bestmask = mask;
      }
    }
    if (best == 0) return in;  // some factor does not fit alone: keep the plain radices
    for (int i = 0; i < idx_n; ++i)
      if (bestmask >> i & 1) {
        used[i] = 1;
        --left;
      }
    groups[ng++] = best;
  }
  if (ng >= in.nr) return in;
  Radices out;
  out.n = in.n;
  out.nr = ng;
  int ns = 1;
  for (int i = 0; i < ng; ++i) {
    out.r[i] = groups[i];
    out.ns[i] = make_fastdiv(ns);
    out.st[i] = make_fastdiv(in.n / groups[i]);
    ns *= groups[i];
  }
  return out;
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
