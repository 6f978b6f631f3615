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

}  // namespace qmx
