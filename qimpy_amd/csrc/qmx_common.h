// Internal helpers shared by the qmx HIP sources (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/qmx.h"

namespace qmx {

void set_error(const std::string& msg);

#define QMX_CHECK_ARG(cond, msg)      \
  do {                                \
    if (!(cond)) {                    \
      qmx::set_error(msg);            \
      return QMX_ERR_ARG;             \
    }                                 \
  } while (0)

#define QMX_HIP(expr)                                                        \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess) {                                                  \
      qmx::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));     \
      return QMX_ERR_HIP;                                                    \
    }                                                                        \
  } while (0)

// Launch-error check after a kernel launch.
#define QMX_LAUNCHED() QMX_HIP(hipGetLastError())

template <typename R>
struct Cplx;
template <>
struct Cplx<float> {
  using T = float2;
};
template <>
struct Cplx<double> {
  using T = double2;
};

template <typename C2>
__device__ __forceinline__ C2 cmul(C2 a, C2 b) {
  return C2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename C2>
__device__ __forceinline__ C2 cadd(C2 a, C2 b) {
  return C2{a.x + b.x, a.y + b.y};
}
template <typename C2>
__device__ __forceinline__ C2 csub(C2 a, C2 b) {
  return C2{a.x - b.x, a.y - b.y};
}
template <typename C2, typename R>
__device__ __forceinline__ C2 cscale(C2 a, R s) {
  return C2{a.x * s, a.y * s};
}
// Multiply by (s * i) where s = +1 or -1.
template <typename C2, int S>
__device__ __forceinline__ C2 mul_si(C2 a) {
  return S > 0 ? C2{-a.y, a.x} : C2{a.y, -a.x};
}

// Exact unsigned division t / d for t * d < 2^32 via a multiply-high with
// magic = ceil(2^32 / d) (d = 1 is special-cased by the caller's magic 0).
struct FastDiv {
  int d;
  uint32_t magic;
};

inline FastDiv make_fastdiv(int d) {
  FastDiv f;
  f.d = d;
  f.magic = d == 1 ? 0u : (uint32_t)((((uint64_t)1 << 32) + (uint64_t)d - 1) / (uint64_t)d);
  return f;
}

__device__ __forceinline__ int fdiv(int t, const FastDiv& f) {
  return f.d == 1 ? t : (int)__umulhi((uint32_t)t, f.magic);
}

constexpr int kMaxRadices = 16;

// Factorisation of an FFT length into the radices 4, 2, 3, 5 (applied in
// that order by the Stockham passes), with the sub-transform length Ns
// entering each pass and its division magic.
struct Radices {
  int n;
  int nr;
  int r[kMaxRadices];
  FastDiv ns[kMaxRadices];  // product of the radices before pass ip
  FastDiv st[kMaxRadices];  // n / r[ip], the butterflies per line of pass ip
};

bool factorize(int n, Radices* out);

}  // namespace qmx
