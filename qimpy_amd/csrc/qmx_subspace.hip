// Small fused kernels of the batched Davidson step (one launch each instead
// of a dozen elementwise tensor operations): residual normalisation with the
// padding constraint, the canonical basis of the residual block, and the
// assembly of the expanded overlap and Hamiltonian (_davidson.py:188-206 with
// TileExpansion's block layout, _davidson.py:240-264).
#include "qmx_common.h"

using namespace qmx;

namespace {

constexpr int kGlueThreads = 256;

// X[k][b][g] *= 1 / norm[k][b] (0 where norm <= 1e-30), zero for g >= n[k].
template <typename C2>
__global__ void __launch_bounds__(kGlueThreads)
    k_normalize_constrain(C2* __restrict__ C, int64_t ldb, int64_t ldk, int nb, int n_tot,
                          const double* __restrict__ norm, const int* __restrict__ nvec) {
  using R = decltype(C2{}.x);
  const int b = blockIdx.x, k = blockIdx.y;
  C2* row = C + k * ldk + (int64_t)b * ldb;
  const double nr = norm[(int64_t)k * nb + b];
  const R s = nr > 1e-30 ? (R)(1.0 / nr) : R(0);
  const int nk_valid = nvec[k];
  for (int g = threadIdx.x; g < n_tot; g += blockDim.x) {
    if (g >= nk_valid) {
      row[g] = C2{0, 0};
    } else {
      const C2 v = row[g];
      row[g] = C2{v.x * s, v.y * s};
    }
  }
}

// Ut[k][j][i] = W[k][i][j] / sqrt(lam[k][j]) where lam[k][j] > tol, else 0
// (the transposed canonical basis; W holds eigenvectors in columns).
template <typename C2>
__global__ void __launch_bounds__(kGlueThreads)
    k_canonical_basis(const double2* __restrict__ W, const double* __restrict__ lam, int m,
                      int64_t total, double tol, C2* __restrict__ Ut) {
  using R = decltype(C2{}.x);
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int64_t mm = (int64_t)m * m;
  const int64_t k = t / mm;
  const int r = (int)(t - k * mm);
  const int j = r / m, i = r - j * m;  // output (j, i)
  const double l = lam[k * m + j];
  const double s = l > tol ? 1.0 / sqrt(l) : 0.0;
  const double2 w = W[k * mm + (int64_t)i * m + j];
  Ut[t] = C2{(R)(w.x * s), (R)(w.y * s)};
}

// Expanded subspace of one Davidson step, n2 = 2 nb, from XO = [C^H R; R^H R]
// and XH = [C^H HR; R^H HR] (each (2 nb) x nb), the current eigenvalues and
// the residual-overlap eigenvalues lam (a residual direction j is dropped
// where lam[j] <= tol; its rows of XO and XH are zero):
//   E = O - I with O = [[I, C^H R], [R^H C, (R^H R + h.c.)/2]], E_jj = 0 on
//       dropped directions (their O_jj is set to 1);
//   H = [[diag(eig), C^H HR], [HR^H C, (R^H HR + h.c.)/2]] + shift on the
//       dropped diagonal.
__global__ void __launch_bounds__(kGlueThreads)
    k_subspace_assemble(const double2* __restrict__ XO, const double2* __restrict__ XH,
                        const double* __restrict__ eig, const double* __restrict__ lam, int nb,
                        int64_t total, double tol, double shift, double2* __restrict__ E,
                        double2* __restrict__ H) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int n2 = 2 * nb;
  const int64_t nn = (int64_t)n2 * n2;
  const int64_t k = t / nn;
  const int r = (int)(t - k * nn);
  const int i = r / n2, j = r - i * n2;
  const double2* xo = XO + k * n2 * nb;
  const double2* xh = XH + k * n2 * nb;
  double2 e, h;
  if (i < nb && j < nb) {
    e = double2{0.0, 0.0};
    h = double2{i == j ? eig[k * nb + i] : 0.0, 0.0};
  } else if (i < nb) {
    e = xo[(int64_t)i * nb + (j - nb)];
    h = xh[(int64_t)i * nb + (j - nb)];
  } else if (j < nb) {
    const double2 a = xo[(int64_t)j * nb + (i - nb)], b = xh[(int64_t)j * nb + (i - nb)];
    e = double2{a.x, -a.y};
    h = double2{b.x, -b.y};
  } else {
    const int ii = i - nb, jj = j - nb;
    const double2 a = xo[(int64_t)i * nb + jj], at = xo[(int64_t)j * nb + ii];
    const double2 b = xh[(int64_t)i * nb + jj], bt = xh[(int64_t)j * nb + ii];
    e = double2{0.5 * (a.x + at.x), 0.5 * (a.y - at.y)};
    h = double2{0.5 * (b.x + bt.x), 0.5 * (b.y - bt.y)};
    if (ii == jj) {
      const bool keep = lam[k * nb + ii] > tol;
      e = keep ? double2{e.x - 1.0, 0.0} : double2{0.0, 0.0};
      h = double2{h.x + (keep ? 0.0 : shift), 0.0};
    }
  }
  E[t] = e;
  H[t] = h;
}

}  // namespace

extern "C" {

int qmx_normalize_constrain(int dtype, void* C, int64_t ldb, int64_t ldk, int64_t nk, int64_t nb,
                            int64_t n_tot, const double* norm, const int32_t* n, void* stream) {
  QMX_CHECK_ARG(C && n && norm, "qmx_normalize_constrain: null argument");
  QMX_CHECK_ARG(nk >= 0 && nk < 65536 && nb >= 0 && nb < (1 << 30) && n_tot >= 0,
                "qmx_normalize_constrain: bad sizes");
  if (nk == 0 || nb == 0) return QMX_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  dim3 grid((unsigned)nb, (unsigned)nk);
  if (dtype == QMX_C64) {
    hipLaunchKernelGGL(k_normalize_constrain<float2>, grid, dim3(kGlueThreads), 0, st,
                       (float2*)C, ldb, ldk, (int)nb, (int)n_tot, norm, n);
  } else if (dtype == QMX_C128) {
    hipLaunchKernelGGL(k_normalize_constrain<double2>, grid, dim3(kGlueThreads), 0, st,
                       (double2*)C, ldb, ldk, (int)nb, (int)n_tot, norm, n);
  } else {
    set_error("qmx_normalize_constrain: dtype must be QMX_C64 or QMX_C128");
    return QMX_ERR_ARG;
  }
  QMX_LAUNCHED();
  return QMX_OK;
}

int qmx_canonical_basis(int dtype, const void* W, const double* lam, int64_t batch, int64_t m,
                        double tol, void* Ut, void* stream) {
  QMX_CHECK_ARG(W && lam && Ut, "qmx_canonical_basis: null argument");
  QMX_CHECK_ARG(batch >= 0 && m >= 0 && m < (1 << 15), "qmx_canonical_basis: bad sizes");
  const int64_t total = batch * m * m;
  if (total == 0) return QMX_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const unsigned grid = (unsigned)((total + kGlueThreads - 1) / kGlueThreads);
  if (dtype == QMX_C64) {
    hipLaunchKernelGGL(k_canonical_basis<float2>, dim3(grid), dim3(kGlueThreads), 0, st,
                       (const double2*)W, lam, (int)m, total, tol, (float2*)Ut);
  } else if (dtype == QMX_C128) {
    hipLaunchKernelGGL(k_canonical_basis<double2>, dim3(grid), dim3(kGlueThreads), 0, st,
                       (const double2*)W, lam, (int)m, total, tol, (double2*)Ut);
  } else {
    set_error("qmx_canonical_basis: dtype must be QMX_C64 or QMX_C128");
    return QMX_ERR_ARG;
  }
  QMX_LAUNCHED();
  return QMX_OK;
}

int qmx_subspace_assemble(const void* XO, const void* XH, const double* eig, const double* lam,
                          int64_t batch, int64_t nb, double tol, double shift, void* E, void* H,
                          void* stream) {
  QMX_CHECK_ARG(XO && XH && eig && lam && E && H, "qmx_subspace_assemble: null argument");
  QMX_CHECK_ARG(batch >= 0 && nb >= 0 && nb < (1 << 14), "qmx_subspace_assemble: bad sizes");
  const int64_t total = batch * 4 * nb * nb;
  if (total == 0) return QMX_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const unsigned grid = (unsigned)((total + kGlueThreads - 1) / kGlueThreads);
  hipLaunchKernelGGL(k_subspace_assemble, dim3(grid), dim3(kGlueThreads), 0, st,
                     (const double2*)XO, (const double2*)XH, eig, lam, (int)nb, total, tol, shift,
                     (double2*)E, (double2*)H);
  QMX_LAUNCHED();
  return QMX_OK;
}

}  // extern "C"
