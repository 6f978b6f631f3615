// Projector generation on the device (Ions.get_projectors,
// dft/ions/_ions_projectors.py:14-72):
//   beta[k][atom * n_proj + p][g] = R_{i_rf[p]}(|G + k|) * (i^l Y_lm)(G + k)
//                                   * exp(-2 pi i (iG + k) . tau_atom) * scale
// with R a quintic blip spline (math/quintic_spline.py:56-128) and Y_lm the
// real solid harmonics (math/spherical_harmonics.py:53-137). One thread per
// plane wave g of k-point blockIdx.y and atom group blockIdx.z: it evaluates
// the radial functions and harmonics of its G + k once into LDS slots of its
// own, then writes every (atom, projector) entry of its atoms, consecutive
// threads on consecutive g. Atom groups fill the chip when the plane waves
// of the k-points alone give few workgroups.
//
// Harmonics: the associated Legendre recurrence in Cartesian form. With
// Q_l^m = r^l P_l^m(cos t) / (r sin t)^m (no Condon-Shortley phase),
//   Q_m^m = (2m - 1)!!,  Q_l^m = ((2l - 1) z Q_{l-1}^m - (l + m - 1) r^2 Q_{l-2}^m) / (l - m),
// and r^l Y_lm = N_lm Q_l^|m| * {Re, 1, Im}((x + i y)^|m|) for m {>, =, <} 0,
// N_lm = sqrt((2l + 1) / (4 pi) (l - |m|)! / (l + |m|)!) (times sqrt 2 for
// m != 0); this is the convention of the reference's tables (real part at
// index l + m, imaginary part at l - m, m > 0, of (-1)^m r^l Y_l^m).
//
// Spline: the blip centred on node c - 2 has coefficient c; on the interval
// [i, i + 1) of t = |G + k| / dG, with f = t - i, coefficient i + j carries
// the weight w_j(f) = (1/66) sum_{q <= 5 - j} (-1)^q C(6, q) (f + 5 - j - q)^5
// (the cardinal quintic B-spline in truncated-power form, scaled by 120/66 so
// that a blip is 1 at its centre), expanded on the host into a polynomial in f.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>

#include "qmx_common.h"

namespace qmx {
void prof_begin(hipStream_t st, hipEvent_t* a);
void prof_end(hipStream_t st, hipEvent_t a, int id, double bytes);
}  // namespace qmx

namespace {

constexpr int kProjThreads = 128;
constexpr int kProjMax = 32;  // projectors per atom of one species (2 per l up to l = 3)
constexpr int kLMax = 6;

struct ProjDesc {
  int n_proj;
  signed char l[kProjMax];
  signed char m[kProjMax];
  signed char i_rf[kProjMax];
  double norm[kProjMax];  // N_lm, sqrt 2 included for m != 0
  double w[6][6];         // w_j(f) = sum_p w[j][p] f^p
  double G[9];            // Cartesian (G + k)_a = sum_b G[3 a + b] (iG + k)_b
  double gauss[kProjMax];  // Gaussian radials: R(G) = exp(gauss[p] G^2)
};

__device__ __forceinline__ double solid_harmonic(int l, int m, double x, double y, double z,
                                                 double r2) {
  const int am = m < 0 ? -m : m;
  double cr = 1.0, ci = 0.0;  // (x + i y)^|m|
  for (int q = 0; q < am; ++q) {
    const double t = cr * x - ci * y;
    ci = cr * y + ci * x;
    cr = t;
  }
  double q1 = 1.0;  // Q_|m|^|m| = (2|m| - 1)!!
  for (int q = 1; q <= am; ++q) q1 *= (double)(2 * q - 1);
  double q0 = 0.0;
  for (int ll = am + 1; ll <= l; ++ll) {
    const double qn = ((2 * ll - 1) * z * q1 - (ll + am - 1) * r2 * q0) / (double)(ll - am);
    q0 = q1;
    q1 = qn;
  }
  return q1 * (m > 0 ? cr : (m < 0 ? ci : 1.0));
}

template <bool GAUSS>
__global__ void __launch_bounds__(kProjThreads)
    k_projectors(const int* __restrict__ iG, const double* __restrict__ kvec,
                 const int* __restrict__ n, const double* __restrict__ pos, int n_atoms, int apg,
                 const double* __restrict__ coeff, int n_coeff, int n_rf, double inv_dG,
                 double scale, ProjDesc d, double2* __restrict__ out, int64_t ldp, int64_t ldk,
                 int n_tot) {
  __shared__ double val[kProjMax][kProjThreads];
  const int tid = threadIdx.x;
  const int g = blockIdx.x * kProjThreads + tid;
  const int k = blockIdx.y;
  const int a0 = blockIdx.z * apg, a1 = min(n_atoms, a0 + apg);
  if (g >= n_tot) return;
  double2* o = out + (int64_t)k * ldk + g;
  const int np = d.n_proj;
  if (g >= n[k]) {  // padding
    for (int a = a0; a < a1; ++a)
      for (int p = 0; p < np; ++p) o[(int64_t)(a * np + p) * ldp] = double2{0.0, 0.0};
    return;
  }
  double f[3];
  for (int c = 0; c < 3; ++c) f[c] = (double)iG[((int64_t)k * n_tot + g) * 3 + c] + kvec[k * 3 + c];
  double Gc[3];
  for (int a = 0; a < 3; ++a) Gc[a] = d.G[3 * a] * f[0] + d.G[3 * a + 1] * f[1] + d.G[3 * a + 2] * f[2];
  const double r2 = Gc[0] * Gc[0] + Gc[1] * Gc[1] + Gc[2] * Gc[2];
  if constexpr (GAUSS) {
    for (int p = 0; p < np; ++p)
      val[p][tid] = scale * d.norm[p] * exp(d.gauss[p] * r2) *
                    solid_harmonic(d.l[p], d.m[p], Gc[0], Gc[1], Gc[2], r2);
  } else {
    const double t = sqrt(r2) * inv_dG;
    int i = (int)floor(t);
    double fr = t - i;
    if (i > n_coeff - 6) {  // beyond the table (the host checks G_max): clamp, stay in bounds
      i = n_coeff - 6;
      fr = 1.0;
    }
    double w[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      double s = d.w[j][5];
#pragma unroll
      for (int q = 4; q >= 0; --q) s = s * fr + d.w[j][q];
      w[j] = s;
    }
    for (int p = 0; p < np; ++p) {
      const double* cf = coeff + (int64_t)i * n_rf + d.i_rf[p];
      double R = 0.0;
#pragma unroll
      for (int j = 0; j < 6; ++j) R += w[j] * cf[(int64_t)j * n_rf];
      val[p][tid] = scale * d.norm[p] * R * solid_harmonic(d.l[p], d.m[p], Gc[0], Gc[1], Gc[2], r2);
    }
  }
  for (int a = a0; a < a1; ++a) {
    const double dot = f[0] * pos[3 * a] + f[1] * pos[3 * a + 1] + f[2] * pos[3 * a + 2];
    double s, c;
    sincospi(-2.0 * dot, &s, &c);  // exp(-2 pi i (iG + k) . tau)
    for (int p = 0; p < np; ++p) {
      const double v = val[p][tid];
      double2 e;
      switch (d.l[p] & 3) {  // i^l (c + i s)
        case 0: e = double2{c, s}; break;
        case 1: e = double2{-s, c}; break;
        case 2: e = double2{-c, -s}; break;
        default: e = double2{s, -c}; break;
      }
      o[(int64_t)(a * np + p) * ldp] = double2{v * e.x, v * e.y};
    }
  }
}

// j_l(x) / x^l for l <= 6: the Taylor series sum_i (-x^2/2)^i / (i! (2l + 2i + 1)!!)
// below x = 1 + 0.7 l (9 + 2l terms), else j_0 = sin x / x,
// j_1 / x = (j_0 - cos x) / x^2 and the upward recurrence
// j_l / x^l = ((2l - 1) j_{l-1} / x^{l-1} - j_{l-2} / x^{l-2}) / x^2.
__device__ double jl_by_xl(int l, double x) {
  if (x <= 1.0 + 0.7 * l) {
    double term = 1.0;
    for (int q = 1; q <= l; ++q) term /= (double)(2 * q + 1);  // 1 / (2l + 1)!!
    double sum = term;
    const double x2 = x * x;
    for (int i = 1; i < 9 + 2 * l; ++i) {
      term *= -0.5 * x2 / (i * (2.0 * l + 2.0 * i + 1.0));
      sum += term;
    }
    return sum;
  }
  double s, c;
  sincos(x, &s, &c);
  const double x2 = x * x;
  double jm2 = s / x;  // j_0
  if (l == 0) return jm2;
  double jm1 = (jm2 - c) / x2;  // j_1 / x
  for (int q = 2; q <= l; ++q) {
    const double j = ((2 * q - 1) * jm1 - jm2) / x2;
    jm2 = jm1;
    jm1 = j;
  }
  return jm1;
}

constexpr int kRadialThreads = 256;
constexpr int kRadialMaxF = 64;

struct LTable {
  signed char l[kRadialMaxF];
};

// f_tilde[i][iG] = sum_r w[r] f[i][r] r^(2 l_i) j_l(G r) / (G r)^l, G = iG dG:
// one workgroup per (block of 256 G, function i); r staged through LDS.
__global__ void __launch_bounds__(kRadialThreads)
    k_radial_transform(const double* __restrict__ r, const double* __restrict__ w,
                       const double* __restrict__ f, LTable lt, int n_r, int nG, double dG,
                       double* __restrict__ out) {
  __shared__ double rs[kRadialThreads], ws[kRadialThreads];
  const int i = blockIdx.y, l = lt.l[i];
  const int iG = blockIdx.x * kRadialThreads + threadIdx.x;
  const double G = iG * dG;
  double acc = 0.0;
  for (int r0 = 0; r0 < n_r; r0 += kRadialThreads) {
    const int ir = r0 + threadIdx.x;
    if (ir < n_r) {
      const double rr = r[ir];
      double r2l = 1.0;
      for (int q = 0; q < l; ++q) r2l *= rr * rr;
      rs[threadIdx.x] = rr;
      ws[threadIdx.x] = w[ir] * f[(int64_t)i * n_r + ir] * r2l;
    }
    __syncthreads();
    const int m = min(kRadialThreads, n_r - r0);
    for (int q = 0; q < m; ++q) acc += ws[q] * jl_by_xl(l, G * rs[q]);
    __syncthreads();
  }
  if (iG < nG) out[(int64_t)i * nG + iG] = acc;
}

double factorial(int q) {
  double r = 1.0;
  for (int i = 2; i <= q; ++i) r *= i;
  return r;
}

// w[j][p]: coefficient of f^p in (1/66) sum_{q <= 5 - j} (-1)^q C(6, q) (f + 5 - j - q)^5.
void blip_polynomials(double w[6][6]) {
  static const double C6[7] = {1, 6, 15, 20, 15, 6, 1};
  static const double C5[6] = {1, 5, 10, 10, 5, 1};
  for (int j = 0; j < 6; ++j) {
    for (int p = 0; p < 6; ++p) w[j][p] = 0.0;
    for (int q = 0; q <= 5 - j; ++q) {
      const double sgn = (q & 1) ? -1.0 : 1.0;
      const double s = 5 - j - q;  // (f + s)^5 = sum_p C(5, p) f^p s^(5 - p)
      for (int p = 0; p < 6; ++p) w[j][p] += sgn * C6[q] * C5[p] * std::pow(s, 5 - p) / 66.0;
    }
  }
}

}  // namespace

// Both projector entry points: spline radials (rc null) or Gaussian ones
// R_i(G) = exp(-rc_i^2 G^2 / 2) (coeff null).
static int projectors_impl(int64_t nk, int64_t n_tot, const int32_t* iG, const double* k,
                           const int32_t* n, const double* Gbasis, int64_t n_atoms,
                           const double* positions, int64_t n_proj, const int32_t* l,
                           const int32_t* m, const int32_t* i_rf, int64_t n_rf, int64_t n_coeff,
                           const double* coeff, double dG, double G_max, const double* rc,
                           double scale, void* out, int64_t ldp, int64_t ldk, void* stream) {
  using qmx::set_error;
  const bool gauss = rc != nullptr;
  if (!iG || !k || !n || !Gbasis || !l || !m || !i_rf || (!gauss && !coeff) || !out ||
      (n_atoms > 0 && !positions)) {
    set_error("qmx_projectors: null argument");
    return QMX_ERR_ARG;
  }
  if (nk < 0 || nk > 65535 || n_tot < 0 || n_tot > (1 << 30) || n_atoms < 0 || n_proj < 0 ||
      n_rf < 1 || n_rf > 127) {
    set_error("qmx_projectors: bad sizes");
    return QMX_ERR_ARG;
  }
  if (n_proj > kProjMax) {
    set_error("qmx_projectors: more than 32 projectors per atom");
    return QMX_ERR_ARG;
  }
  if (!gauss &&
      (!(dG > 0.0) || !(G_max >= 0.0) || n_coeff < 6 || std::floor(G_max / dG) + 5 >= n_coeff)) {
    set_error("qmx_projectors: the radial table does not reach G_max");
    return QMX_ERR_ARG;
  }
  if (ldp < n_tot || (nk > 1 && ldk < n_atoms * n_proj * ldp)) {
    set_error("qmx_projectors: bad strides");
    return QMX_ERR_ARG;
  }
  ProjDesc d{};
  d.n_proj = (int)n_proj;
  for (int p = 0; p < n_proj; ++p) {
    const int lp = l[p], mp = m[p];
    if (lp < 0 || lp > kLMax || mp < -lp || mp > lp || i_rf[p] < 0 || i_rf[p] >= n_rf) {
      set_error("qmx_projectors: bad (l, m, i_rf) of a projector");
      return QMX_ERR_ARG;
    }
    const int am = mp < 0 ? -mp : mp;
    d.l[p] = (signed char)lp;
    d.m[p] = (signed char)mp;
    d.i_rf[p] = (signed char)i_rf[p];
    d.norm[p] = std::sqrt((2 * lp + 1) / (4.0 * M_PI) * factorial(lp - am) / factorial(lp + am)) *
                (mp != 0 ? std::sqrt(2.0) : 1.0);
    if (gauss) {
      const double r = rc[i_rf[p]];
      if (!(r >= 0.0) || !std::isfinite(r)) {
        set_error("qmx_gaussian_projectors: bad Gaussian width");
        return QMX_ERR_ARG;
      }
      d.gauss[p] = -0.5 * r * r;
    }
  }
  blip_polynomials(d.w);
  for (int i = 0; i < 9; ++i) d.G[i] = Gbasis[i];
  if (nk == 0 || n_tot == 0 || n_atoms == 0 || n_proj == 0) return QMX_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int64_t gx = (n_tot + kProjThreads - 1) / kProjThreads;
  int64_t gz = std::min<int64_t>(n_atoms, std::max<int64_t>(1, (1024 + gx * nk - 1) / (gx * nk)));
  const int64_t apg = (n_atoms + gz - 1) / gz;  // atoms per group
  gz = (n_atoms + apg - 1) / apg;
  dim3 grid((unsigned)gx, (unsigned)nk, (unsigned)gz);
  hipEvent_t ev;
  qmx::prof_begin(st, &ev);
  if (gauss)
    hipLaunchKernelGGL(k_projectors<true>, grid, dim3(kProjThreads), 0, st, iG, k, n, positions,
                       (int)n_atoms, (int)apg, nullptr, 0, (int)n_rf, 0.0, scale, d,
                       static_cast<double2*>(out), ldp, ldk, (int)n_tot);
  else
    hipLaunchKernelGGL(k_projectors<false>, grid, dim3(kProjThreads), 0, st, iG, k, n, positions,
                       (int)n_atoms, (int)apg, coeff, (int)n_coeff, (int)n_rf, 1.0 / dG, scale, d,
                       static_cast<double2*>(out), ldp, ldk, (int)n_tot);
  QMX_LAUNCHED();
  // compulsory bytes: the complex128 output
  qmx::prof_end(st, ev, QMX_KERNEL_PROJ + QMX_KERNEL_C128, 16.0 * nk * n_atoms * n_proj * n_tot);
  return QMX_OK;
}

extern "C" int qmx_projectors(int64_t nk, int64_t n_tot, const int32_t* iG, const double* k,
                              const int32_t* n, const double* Gbasis, int64_t n_atoms,
                              const double* positions, int64_t n_proj, const int32_t* l,
                              const int32_t* m, const int32_t* i_rf, int64_t n_rf,
                              int64_t n_coeff, const double* coeff, double dG, double G_max,
                              double scale, void* out, int64_t ldp, int64_t ldk, void* stream) {
  if (!coeff) {
    qmx::set_error("qmx_projectors: null argument");
    return QMX_ERR_ARG;
  }
  return projectors_impl(nk, n_tot, iG, k, n, Gbasis, n_atoms, positions, n_proj, l, m, i_rf,
                         n_rf, n_coeff, coeff, dG, G_max, nullptr, scale, out, ldp, ldk, stream);
}

extern "C" int qmx_gaussian_projectors(int64_t nk, int64_t n_tot, const int32_t* iG,
                                       const double* k, const int32_t* n, const double* Gbasis,
                                       int64_t n_atoms, const double* positions, int64_t n_proj,
                                       const int32_t* l, const int32_t* m, const int32_t* i_rf,
                                       int64_t n_rf, const double* rc, double scale, void* out,
                                       int64_t ldp, int64_t ldk, void* stream) {
  if (!rc) {
    qmx::set_error("qmx_gaussian_projectors: null argument");
    return QMX_ERR_ARG;
  }
  return projectors_impl(nk, n_tot, iG, k, n, Gbasis, n_atoms, positions, n_proj, l, m, i_rf,
                         n_rf, 0, nullptr, 0.0, 0.0, rc, scale, out, ldp, ldk, stream);
}

extern "C" int qmx_radial_transform(int64_t n_f, int64_t n_r, const double* r, const double* w,
                                    const double* f, const int32_t* l, int64_t nG, double dG,
                                    double* out, void* stream) {
  using qmx::set_error;
  if (!r || !w || !f || !l || !out) {
    set_error("qmx_radial_transform: null argument");
    return QMX_ERR_ARG;
  }
  if (n_f < 0 || n_f > kRadialMaxF || n_r < 0 || n_r > (1 << 30) || nG < 0 ||
      nG > (1 << 30) || !(dG > 0.0)) {
    set_error("qmx_radial_transform: bad sizes (at most 64 functions)");
    return QMX_ERR_ARG;
  }
  LTable lt{};
  for (int i = 0; i < n_f; ++i) {
    if (l[i] < 0 || l[i] > kLMax) {
      set_error("qmx_radial_transform: l must be in 0..6");
      return QMX_ERR_ARG;
    }
    lt.l[i] = (signed char)l[i];
  }
  if (n_f == 0 || nG == 0) return QMX_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  dim3 grid((unsigned)((nG + kRadialThreads - 1) / kRadialThreads), (unsigned)n_f);
  hipLaunchKernelGGL(k_radial_transform, grid, dim3(kRadialThreads), 0, st, r, w, f, lt,
                     (int)n_r, (int)nG, dG, out);
  QMX_LAUNCHED();
  return QMX_OK;
}
