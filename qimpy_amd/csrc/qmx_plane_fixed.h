// The plane pass of the local apply specialised at compile time for square
// planes of the batched and served workloads (qmx_plane_fixed.hip).
#pragma once

#include "qmx_common.h"

namespace qmx {

// True when k_plane_fixed exists for an nx x ny plane with ncol active
// columns (nx == ny in 24, 30, 32, 36, 40, 48, 60, 64).
bool plane_fixed_available(int nx, int ny, int ncol);

// Launch it: grid (nz, bands, k-points), lds bytes as for k_plane_potential
// (the plane with pitch ny | 1, both twiddle tables, the active-row list).
int launch_plane_fixed(bool c128, int n, dim3 grid, size_t lds, hipStream_t st, void* T1,
                       const void* Vt, const int* vslot, const int* col_ix, const int* col_iy,
                       const int* x_list, const void* twx, const void* twy, int ncol, int nxa,
                       int nb);

// For a plane size whose fixed kernel runs its transforms in prime-factor
// form (n = A B, gcd(A, B) = 1): A and B, true. The kernel then wants each
// potential box with x and y in the rho order: value (x, y) at
// x' ny + y', n' = B ((n (B^-1 mod A)) mod A) + (n (A^-1 mod B)) mod B.
bool plane_fixed_space_order(int n, int* A, int* B);

// The column passes with the z length fixed at compile time
// (qmx_cols_fixed.hip): one thread per active column, kColFixedThreads
// columns per workgroup, grid (column blocks, bands, k-points).
constexpr int kColFixedThreads = 128;

struct ColFixedArgs {
  dim3 grid;
  hipStream_t st;
  const void* C;  // sphere input (first pass) or the KE operand (last pass)
  int64_t ldb, ldk;
  void* T1;  // z-major columns, T1[k][band][z][col]
  void* out;  // sphere output of the last pass
  int64_t oldb, oldk;
  const void* ke;  // kinetic energies (last pass; null: no KE term)
  int64_t ke_ldk;
  const int* cstart;  // [k][ncol + 1]
  const int* zrange;  // [k][ncol] packed (lo, hi) z interval of each column's points
  const int* nvec;    // [k] basis size (last pass zeroes the padding past it)
  const void* tw;     // exact z twiddle table
  int ncol, nb, n_tot, b0;
  int run_max;  // most sphere points over kColFixedThreads consecutive columns
  double scale;
  int accumulate;
};

// True when the fixed column passes exist for this z length and precision.
bool cols_fixed_available(int nz, bool c128);

// First pass (to_sphere false: sphere -> T1, inverse z) or last pass
// (to_sphere true: T1 -> sphere, forward z, scale, + KE C, padding zeroed).
int launch_cols_fixed(bool c128, bool to_sphere, int nz, const ColFixedArgs& a);

}  // namespace qmx
