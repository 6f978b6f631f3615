// Sphere-sparse local-potential apply and dense 3D FFT for gfx950.
//
// The plane-wave sphere occupies ~1/15 of the FFT box (4x Ecut density
// box), so the apply never materialises a full box. With the box (n0, n1,
// n2) C-ordered (x, y, z; z fastest), one band goes through five LDS-staged
// passes, each a batch of 1D Stockham transforms:
//   A  sphere -> T1[col][z]      scatter in-sphere points of each active
//                                (x, y) column, inverse z-FFT
//   B  T1 -> T2[xa][y][z]        inverse y-FFT on active x-planes only
//   C  T2 -> T2 (in place)       inverse x-FFT, x V(r), forward x-FFT,
//                                keep the active x-planes only
//   D  T2 -> T1                  forward y-FFT, keep the active columns
//   E  T1 -> sphere              forward z-FFT, gather the sphere, x 1/N,
//                                + KE C, zero the padding
// Intermediates are read and written coalesced: z is the unit-stride axis
// of every tile. Bands are processed in blocks sized to the workspace.
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "qmx_common.h"
#include "qmx_tune.h"
#include <algorithm>

#include "qmx_fft_core.h"
#include "qmx_plane_fixed.h"

namespace qmx {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

// ------------------------------------------------------------ profiling
// When enabled, every kernel launch of the library is bracketed by a pair of
// HIP events on its own stream and tagged with a kernel id and the
// compulsory bytes it moves; qmx_profile_read() resolves the events.
struct ProfRec {
  int id;
  hipEvent_t a, b;
  double bytes, flops;
};
static std::mutex g_prof_mu;
static bool g_prof_on = false;
static std::vector<ProfRec> g_prof;
// Resolved events go back to a free list per device (creating an event is a
// driver call far slower than recording one), so a profiled stream adds
// little host time per launch.
static std::vector<std::pair<int, hipEvent_t>> g_prof_free;

bool prof_on() { return g_prof_on; }

static hipEvent_t prof_event() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  {
    std::lock_guard<std::mutex> lock(g_prof_mu);
    for (size_t i = g_prof_free.size(); i-- > 0;) {
      if (g_prof_free[i].first == dev) {
        hipEvent_t e = g_prof_free[i].second;
        g_prof_free[i] = g_prof_free.back();
        g_prof_free.pop_back();
        return e;
      }
    }
  }
  hipEvent_t e = nullptr;
  return hipEventCreate(&e) == hipSuccess ? e : nullptr;
}

static void prof_release(hipEvent_t e) {
  int dev = 0;
  if (!e) return;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipEventDestroy(e);
    return;
  }
  std::lock_guard<std::mutex> lock(g_prof_mu);
  g_prof_free.push_back({dev, e});
}

void prof_begin(hipStream_t st, hipEvent_t* a) {
  *a = nullptr;
  if (!g_prof_on) return;
  *a = prof_event();
  if (*a && hipEventRecord(*a, st) != hipSuccess) {
    prof_release(*a);
    *a = nullptr;
  }
}

void prof_end_flops(hipStream_t st, hipEvent_t a, int id, double bytes, double flops) {
  if (!a) return;
  hipEvent_t b = prof_event();
  if (!b || hipEventRecord(b, st) != hipSuccess) {
    prof_release(a);
    prof_release(b);
    return;
  }
  std::lock_guard<std::mutex> lock(g_prof_mu);
  g_prof.push_back({id, a, b, bytes, flops});
}

void prof_end(hipStream_t st, hipEvent_t a, int id, double bytes) {
  prof_end_flops(st, a, id, bytes, 0.0);
}

bool factorize(int n, Radices* out) {
  if (n < 1) return false;
  out->n = n;
  out->nr = 0;
  int m = n, ns = 1;
  const int order[4] = {4, 2, 3, 5};
  for (int r : order) {
    while (m % r == 0) {
      if (out->nr >= kMaxRadices) return false;
      out->r[out->nr] = r;
      out->ns[out->nr] = make_fastdiv(ns);
      out->st[out->nr] = make_fastdiv(n / r);
      ++out->nr;
      ns *= r;
      m /= r;
    }
  }
  return m == 1;
}

constexpr int kThreads = 256;
constexpr int kLdsBudget = 60 * 1024;
constexpr int kLdsMax = 160 * 1024;  // LDS of one workgroup on gfx950

// Largest line count M for lines of length n with two ping-pong buffers,
// the twiddle table and `extra_per_line` bytes per line within the budget.
static int max_lines(int n, int csize, int extra_per_line, int cap) {
  int m = cap;
  while (m > 1 && (2 * (m + 1) * n * csize + n * csize + m * extra_per_line) > kLdsBudget) --m;
  return m;
}

// Split `extent` into the fewest tiles of at most `mmax` and return the tile.
static int even_tile(int extent, int mmax) {
  int ntile = (extent + mmax - 1) / mmax;
  return (extent + ntile - 1) / ntile;
}

__host__ __device__ inline int odd_ld(int m) { return m | 1; }

}  // namespace qmx

using namespace qmx;

struct qmx_plan {
  int device;
  int shape[3];
  int nk, n_tot;
  int64_t N;
  int nxa, ncol;
  int64_t n_sum;  // sum of n[k]
  Radices rad[3];
  int* d_x_list;      // [nxa] box x of each active plane
  int* d_pcol_start;  // [nxa + 1] first column of each active plane
  int* d_col_iy;      // [ncol] y of each column
  int* d_cstart;      // [nk][ncol + 1] first sphere point of each column
  int* d_t1idx;       // [nk][n_tot] col * n2 + z of each sphere point
  int* d_n;           // [nk]
  int* d_vslot;       // [nk] potential box of each k
  int* d_col_ix;      // [ncol] x of each column
  int* d_zrange;      // [nk][ncol] z interval of each column's points, packed (lo, hi)
  bool col_interval;  // every column's points form one cyclic z interval
  int col_run_max;    // most points of kColFixedThreads consecutive columns of one k
  int local_path;     // 0 automatic, 1 separate passes B, C, D
  int64_t* h_nsum;    // [nk + 1] prefix sums of n[k] (host)
  int n_vslots;       // 1 + largest slot
  float2* d_tw32[3];
  double2* d_tw64[3];
  void* d_block;  // device block holding the integer tables
  void* h_stage;  // pinned staging copy of the block
  size_t block_bytes;    // capacity of d_block and h_stage
  hipEvent_t uploaded;   // recorded after the copy out of h_stage
  // Stream ordering of the plan's users (plan_enter / plan_leave): the
  // event after the latest call that read the tables, and its stream.
  hipEvent_t last_use;
  hipStream_t last_stream;
  bool entered;
};

namespace {

// ---------------------------------------------------------------- pass A
// One workgroup: M consecutive columns of one band of one k.
template <typename C2, typename R, bool ZMAJOR>
__global__ void __launch_bounds__(kThreads)
    k_sphere_to_cols(const C2* __restrict__ C, int64_t ldb, int64_t ldk,
                     C2* __restrict__ T1, const int* __restrict__ t1idx,
                     const int* __restrict__ cstart, const C2* __restrict__ tw_g, Radices rz,
                     FastDiv fm, int ncol, int nb, int n_tot, int b0) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int nz = rz.n, M = fm.d, ld = odd_ld(M);
  C2* tw = reinterpret_cast<C2*>(smem_raw);
  C2* a = tw + nz;
  C2* b = a + nz * ld;
  const int col0 = blockIdx.x * M, ib = blockIdx.y, k = blockIdx.z;
  const int mcols = min(M, ncol - col0);
  load_twiddles(tw, tw_g, nz);
  zero_lds(a, nz * ld);
  __syncthreads();
  const int* cs = cstart + (int64_t)k * (ncol + 1);
  const int gs = cs[col0], ge = cs[col0 + mcols];
  const C2* Cb = C + k * ldk + (int64_t)(b0 + ib) * ldb;
  const int* ti = t1idx + (int64_t)k * n_tot;
  const int base = col0 * nz;
  for (int g = gs + threadIdx.x; g < ge; g += blockDim.x) {
    const int t = ti[g] - base;
    const int l = t / nz;
    a[(t - l * nz) * ld + l] = Cb[g];
  }
  __syncthreads();
  C2* res = fft_lds<C2, +1>(a, b, tw, rz, fm, ld);
  C2* slab = T1 + ((int64_t)k * nb + ib) * ncol * nz;
  if (ZMAJOR) {  // T1[z][col]: runs of mcols columns per z
    for (int i = threadIdx.x; i < M * nz; i += blockDim.x) {
      const int p = fdiv(i, fm);
      const int l = i - p * M;
      if (l < mcols) slab[(int64_t)p * ncol + col0 + l] = res[p * ld + l];
    }
  } else {  // T1[col][z]
    C2* out = slab + (int64_t)col0 * nz;
    for (int i = threadIdx.x; i < mcols * nz; i += blockDim.x) {
      const int l = i / nz;
      const int p = i - l * nz;
      out[i] = res[p * ld + l];
    }
  }
}

// ---------------------------------------------------------------- pass B
// One workgroup: one active x-plane, one z-tile, one (k, band).
template <typename C2, typename R>
__global__ void __launch_bounds__(kThreads)
    k_cols_to_planes(const C2* __restrict__ T1, C2* __restrict__ T2,
                     const int* __restrict__ pcol_start, const int* __restrict__ col_iy,
                     const C2* __restrict__ tw_g, Radices ry, FastDiv fm, int ncol, int nxa,
                     int nz) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int ny = ry.n, M = fm.d, ld = odd_ld(M);
  C2* tw = reinterpret_cast<C2*>(smem_raw);
  C2* a = tw + ny;
  C2* b = a + ny * ld;
  const int z0 = blockIdx.x * M, xa = blockIdx.y;
  const int64_t bk = blockIdx.z;
  const int mz = min(M, nz - z0);
  load_twiddles(tw, tw_g, ny);
  zero_lds(a, ny * ld);
  __syncthreads();
  const int c0 = pcol_start[xa], nc = pcol_start[xa + 1] - c0;
  const C2* src = T1 + (bk * ncol + c0) * nz + z0;
  for (int i = threadIdx.x; i < nc * M; i += blockDim.x) {
    const int c = fdiv(i, fm);
    const int l = i - c * M;
    if (l < mz) a[col_iy[c0 + c] * ld + l] = src[(int64_t)c * nz + l];
  }
  __syncthreads();
  C2* res = fft_lds<C2, +1>(a, b, tw, ry, fm, ld);
  C2* dst = T2 + (bk * nxa + xa) * (int64_t)ny * nz + z0;
  for (int i = threadIdx.x; i < ny * M; i += blockDim.x) {
    const int y = fdiv(i, fm);
    const int l = i - y * M;
    if (l < mz) dst[(int64_t)y * nz + l] = res[y * ld + l];
  }
}

// ---------------------------------------------------------------- pass C
// One workgroup: one y, one z-tile of Mz, one k, a chunk of bands. Lines are
// (band, z) pairs, nbg bands at a time (M = nbg * Mz lines, M / 4 per wave);
// each wave transforms its own lines with wave-level LDS fences only, so a
// group costs three workgroup barriers. V(x, y, z-tile) is staged once.
template <typename C2, typename R>
__global__ void __launch_bounds__(kThreads)
    k_xpass_potential(C2* __restrict__ T2, const R* __restrict__ V,
                      const int* __restrict__ vslot, const int* __restrict__ x_list,
                      const C2* __restrict__ tw_g, Radices rx, FastDiv fmw, FastDiv fM,
                      int zshift, int nbg, int ny, int nz, int nxa, int nb,
                      int bands_per_chunk, int n_chunks) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int nx = rx.n, M = fM.d, ld = odd_ld(M), Mz = 1 << zshift, zmask = Mz - 1;
  const int mw = fmw.d;
  C2* tw = reinterpret_cast<C2*>(smem_raw);
  C2* a = tw + nx;
  C2* b = a + nx * ld;
  R* vt = reinterpret_cast<R*>(b + nx * ld);
  const int z0 = blockIdx.x * Mz, y = blockIdx.y;
  const int k = blockIdx.z / n_chunks, chunk = blockIdx.z - k * n_chunks;
  const int mz = min(Mz, nz - z0);
  const int bstart = chunk * bands_per_chunk;
  const int bstop = min(nb, bstart + bands_per_chunk);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  load_twiddles(tw, tw_g, nx);
  const R* Vk = V + (int64_t)vslot[k] * nx * ny * nz;
  for (int i = threadIdx.x; i < nx * Mz; i += blockDim.x) {
    const int x = i >> zshift;
    const int zl = i & zmask;
    vt[i] = zl < mz ? Vk[((int64_t)x * ny + y) * nz + z0 + zl] : R(0);
  }
  const int64_t plane = (int64_t)ny * nz;
  for (int g0 = bstart; g0 < bstop; g0 += nbg) {
    __syncthreads();  // previous group's stores have read LDS
    zero_lds(a, nx * ld);
    __syncthreads();
    for (int i = threadIdx.x; i < nxa * M; i += blockDim.x) {
      const int xa = fdiv(i, fM);
      const int l = i - xa * M;
      const int ib = g0 + (l >> zshift), zl = l & zmask;
      if (ib < bstop && zl < mz)
        a[x_list[xa] * ld + l] =
            T2[(((int64_t)k * nb + ib) * nxa + xa) * plane + (int64_t)y * nz + z0 + zl];
    }
    __syncthreads();
    const int l0 = wave * mw;
    C2* r1 = fft_lds_wave<C2, +1>(a, b, tw, rx, fmw, ld, l0);
    C2* other = (r1 == a) ? b : a;
    for (int t = lane; t < nx * mw; t += 64) {
      const int x = fdiv(t, fmw);
      const int l = l0 + t - x * mw;
      r1[x * ld + l] = cscale(r1[x * ld + l], vt[x * Mz + (l & zmask)]);
    }
    wave_lds_sync();
    C2* r2 = fft_lds_wave<C2, -1>(r1, other, tw, rx, fmw, ld, l0);
    __syncthreads();
    for (int i = threadIdx.x; i < nxa * M; i += blockDim.x) {
      const int xa = fdiv(i, fM);
      const int l = i - xa * M;
      const int ib = g0 + (l >> zshift), zl = l & zmask;
      if (ib < bstop && zl < mz)
        T2[(((int64_t)k * nb + ib) * nxa + xa) * plane + (int64_t)y * nz + z0 + zl] =
            r2[x_list[xa] * ld + l];
    }
  }
}

// ----------------------------------------------------- density, pass C
// Density on the separate passes (x-y planes too large for the plane pass):
// after the inverse z- and y-transforms (passes A and B into T2), one
// workgroup takes one y, one z-tile of Mz (a power of two), one k and a chunk
// of bands, nbg bands per group (M = nbg * Mz lines, M / 4 per wave). The
// active x rows of the group's (band, z) lines are scattered into the zeroed
// plane and transformed along x (inverse), each wave its own lines; then the
// thread that owns a point (x, z) adds w |psi|^2 of the group's bands in
// band order to its FP64 accumulator in LDS. After the last group the
// accumulators are added atomically to the x-major density rhoX[x][y][z].
// CROSS: T2b holds the second spinor component of the same bands; both are
// transformed and w psi_1 conj(psi_0) is accumulated, the imaginary part
// into rhoX[N + ...].
template <typename C2, typename R, bool CROSS>
__global__ void __launch_bounds__(kThreads)
    k_xpass_density(const C2* __restrict__ T2, const C2* __restrict__ T2b,
                    const double* __restrict__ prefac, int64_t ldp, const int* __restrict__ x_list,
                    const C2* __restrict__ tw_g, Radices rx, FastDiv fmw, FastDiv fM, int zshift,
                    int nbg, int ny, int nz, int nxa, int nb, int b0, int bands_per_chunk,
                    int n_chunks, double* __restrict__ rhoX) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int kNp = CROSS ? 2 : 1;
  const int nx = rx.n, M = fM.d, ld = odd_ld(M), Mz = 1 << zshift, zmask = Mz - 1;
  const int mw = fmw.d;
  C2* tw = reinterpret_cast<C2*>(smem_raw);
  C2* pa = tw + nx;  // kNp pairs of planes (a, b), nx x ld each
  double* acc = reinterpret_cast<double*>(pa + 2 * kNp * nx * ld);  // kNp x nx x Mz
  const int z0 = blockIdx.x * Mz, y = blockIdx.y;
  const int k = blockIdx.z / n_chunks, chunk = blockIdx.z - k * n_chunks;
  const int mz = min(Mz, nz - z0);
  const int bstart = chunk * bands_per_chunk;
  const int bstop = min(nb, bstart + bands_per_chunk);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int npt = nx * Mz;
  load_twiddles(tw, tw_g, nx);
  for (int t = threadIdx.x; t < kNp * npt; t += blockDim.x) acc[t] = 0.0;
  const int64_t plane = (int64_t)ny * nz;
  const C2* src[2] = {T2, T2b};
  for (int g0 = bstart; g0 < bstop; g0 += nbg) {
    __syncthreads();  // the previous group's planes have been read
    zero_lds(pa, 2 * kNp * nx * ld);
    __syncthreads();
    for (int c = 0; c < kNp; ++c) {
      C2* a = pa + 2 * c * nx * ld;
      for (int i = threadIdx.x; i < nxa * M; i += blockDim.x) {
        const int xa = fdiv(i, fM);
        const int l = i - xa * M;
        const int ib = g0 + (l >> zshift), zl = l & zmask;
        if (ib < bstop && zl < mz)
          a[x_list[xa] * ld + l] =
              src[c][(((int64_t)k * nb + ib) * nxa + xa) * plane + (int64_t)y * nz + z0 + zl];
      }
    }
    __syncthreads();
    const int l0 = wave * mw;
    C2* r[2];
    for (int c = 0; c < kNp; ++c) {
      C2* a = pa + 2 * c * nx * ld;
      r[c] = fft_lds_wave<C2, +1>(a, a + nx * ld, tw, rx, fmw, ld, l0);
    }
    __syncthreads();  // every wave's lines are transformed
    const int ng = min(nbg, bstop - g0);
    for (int t = threadIdx.x; t < npt; t += blockDim.x) {
      const int x = t >> zshift, zl = t & zmask;
      double sr = 0.0, si = 0.0;
      for (int j = 0; j < ng; ++j) {
        const double w = prefac[(int64_t)k * ldp + b0 + g0 + j];
        const C2 v = r[0][x * ld + (j << zshift) + zl];
        if constexpr (CROSS) {  // psi_1 conj(psi_0), psi_0 = v, psi_1 = u
          const C2 u = r[1][x * ld + (j << zshift) + zl];
          sr += w * ((double)u.x * (double)v.x + (double)u.y * (double)v.y);
          si += w * ((double)u.y * (double)v.x - (double)u.x * (double)v.y);
        } else {
          sr += w * ((double)v.x * (double)v.x + (double)v.y * (double)v.y);
        }
      }
      acc[t] += sr;
      if constexpr (CROSS) acc[npt + t] += si;
    }
  }
  __syncthreads();
  const int64_t N = (int64_t)nx * plane;
  for (int t = threadIdx.x; t < npt; t += blockDim.x) {
    const int x = t >> zshift, zl = t & zmask;
    if (zl >= mz) continue;
    const int64_t o = ((int64_t)x * ny + y) * nz + z0 + zl;
    if (acc[t] != 0.0) unsafeAtomicAdd(rhoX + o, acc[t]);
    if constexpr (CROSS)
      if (acc[npt + t] != 0.0) unsafeAtomicAdd(rhoX + N + o, acc[npt + t]);
  }
}

// dst (= or +=) src, n doubles.
template <bool kAccumulate>
__global__ void __launch_bounds__(256)
    k_density_add(const double* __restrict__ src, double* __restrict__ dst, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    if constexpr (kAccumulate) dst[i] += src[i];
    else dst[i] = src[i];
  }
}

// ---------------------------------------------------------------- pass D
template <typename C2, typename R>
__global__ void __launch_bounds__(kThreads)
    k_planes_to_cols(const C2* __restrict__ T2, C2* __restrict__ T1,
                     const int* __restrict__ pcol_start, const int* __restrict__ col_iy,
                     const C2* __restrict__ tw_g, Radices ry, FastDiv fm, int ncol, int nxa,
                     int nz) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int ny = ry.n, M = fm.d, ld = odd_ld(M);
  C2* tw = reinterpret_cast<C2*>(smem_raw);
  C2* a = tw + ny;
  C2* b = a + ny * ld;
  const int z0 = blockIdx.x * M, xa = blockIdx.y;
  const int64_t bk = blockIdx.z;
  const int mz = min(M, nz - z0);
  load_twiddles(tw, tw_g, ny);
  const C2* src = T2 + (bk * nxa + xa) * (int64_t)ny * nz + z0;
  for (int i = threadIdx.x; i < ny * M; i += blockDim.x) {
    const int y = fdiv(i, fm);
    const int l = i - y * M;
    a[y * ld + l] = l < mz ? src[(int64_t)y * nz + l] : C2{0, 0};
  }
  __syncthreads();
  C2* res = fft_lds<C2, -1>(a, b, tw, ry, fm, ld);
  const int c0 = pcol_start[xa], nc = pcol_start[xa + 1] - c0;
  C2* dst = T1 + (bk * ncol + c0) * nz + z0;
  for (int i = threadIdx.x; i < nc * M; i += blockDim.x) {
    const int c = fdiv(i, fm);
    const int l = i - c * M;
    if (l < mz) dst[(int64_t)c * nz + l] = res[col_iy[c0 + c] * ld + l];
  }
}

// ------------------------------------------------------ plane pass (B-C-D)
// When an x-y plane fits in LDS, passes B, C and D fuse into one kernel over
// T1 in z-major layout (T1[k][band][z][col]). One workgroup: the plane z of
// one band of one k, row stride odd_ld(ny). The active columns at this z
// are scattered into the zeroed plane, then in place: inverse y-FFT along
// the active x rows, inverse x-FFT along every y, x V(x, y, z) (V given
// z-major, one box per potential slot), forward x-FFT, forward y-FFT along
// the active rows; the active columns are gathered back to the same place.
// HBM traffic per band is T1 in and out (V stays cached across bands).
constexpr int kPlaneThreads = 512;  // planes above kPlaneSmallPts points
constexpr int kPlaneThreadsSmall = 256;
constexpr int kPlaneSmallPts = 4096;  // nx * ny up to a 64 x 64 plane
constexpr int kPlaneMaxE = 24;       // complex values per thread in one FFT pass (512 threads)
constexpr int kPlaneMaxESmall = 16;  // the same with 256 threads (planes up to 64 x 64)
constexpr int kPlaneColPer = 4;      // active columns staged per thread
constexpr int kPlaneVPer = 16;       // potential values prefetched per thread

// Every global read of a plane is issued before the work that needs it: the
// twiddles, the active columns (kept in registers across the LDS zeroing)
// and the potential of the plane (in registers across the inverse y and x
// transforms), so their latency overlaps LDS work instead of stalling it.
template <typename C2, typename R, int NT, int MAXE>
__global__ void __launch_bounds__(NT)
    k_plane_potential(C2* __restrict__ T1, const R* __restrict__ Vt,
                      const int* __restrict__ vslot, const int* __restrict__ col_ix,
                      const int* __restrict__ col_iy, const int* __restrict__ x_list,
                      const C2* __restrict__ twx_g, const C2* __restrict__ twy_g, Radices rx,
                      Radices ry, FastDiv fny, FastDiv fnxa, int ncol, int nxa, int nb) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int nx = rx.n, ny = ry.n, ldp = odd_ld(ny), nz = gridDim.x;
  C2* P = reinterpret_cast<C2*>(smem_raw);
  C2* twx = P + nx * ldp;
  C2* twy = twx + nx;
  int* xl = reinterpret_cast<int*>(twy + ny);
  const int z = blockIdx.x, ib = blockIdx.y, k = blockIdx.z;
  const int tid = threadIdx.x;
  C2* row = T1 + (((int64_t)k * nb + ib) * nz + z) * ncol;
  // Staged columns as separate real and imaginary parts (an array of
  // double2 held across the zero fill is kept in scratch memory).
  R cvx[kPlaneColPer], cvy[kPlaneColPer];
  int co[kPlaneColPer];
#pragma unroll
  for (int q = 0; q < kPlaneColPer; ++q) {
    const int c = tid + q * NT;
    if (c < ncol) {
      const C2 t = row[c];
      cvx[q] = t.x;
      cvy[q] = t.y;
      co[q] = col_ix[c] * ldp + col_iy[c];
    }
  }
  for (int i = tid; i < nx; i += NT) twx[i] = twx_g[i];
  for (int i = tid; i < ny; i += NT) twy[i] = twy_g[i];
  for (int i = tid; i < nxa; i += NT) xl[i] = x_list[i];
  zero_lds(P, nx * ldp);
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kPlaneColPer; ++q)
    if (tid + q * NT < ncol) P[co[q]] = C2{cvx[q], cvy[q]};
  __syncthreads();
  fft_ip<C2, +1, NT, MAXE, false, true>(P, twy, ry, nxa, fnxa, 1, ldp, xl);
  const R* Vz = Vt + ((int64_t)vslot[k] * nz + z) * nx * ny;
  const int nxy = nx * ny;
  constexpr int kVPer = sizeof(R) == 4 ? kPlaneVPer : kPlaneVPer / 2;
  R vv[kVPer];
#pragma unroll
  for (int q = 0; q < kVPer; ++q) {
    const int i = tid + q * NT;
    vv[q] = i < nxy ? Vz[i] : R(0);
  }
  fft_ip<C2, +1, NT, MAXE, true, false>(P, twx, rx, ny, fny, ldp, 1, nullptr);
#pragma unroll
  for (int q = 0; q < kVPer; ++q) {
    const int i = tid + q * NT;
    if (i < nxy) {
      const int x = fdiv(i, fny);
      const int y = i - x * ny;
      P[x * ldp + y] = cscale(P[x * ldp + y], vv[q]);
    }
  }
  for (int i = tid + kVPer * NT; i < nxy; i += NT) {  // points beyond the prefetch
    const int x = fdiv(i, fny);
    const int y = i - x * ny;
    P[x * ldp + y] = cscale(P[x * ldp + y], Vz[i]);
  }
  __syncthreads();
  fft_ip<C2, -1, NT, MAXE, true, false>(P, twx, rx, ny, fny, ldp, 1, nullptr);
  fft_ip<C2, -1, NT, MAXE, false, true>(P, twy, ry, nxa, fnxa, 1, ldp, xl);
#pragma unroll
  for (int q = 0; q < kPlaneColPer; ++q)
    if (tid + q * NT < ncol) row[tid + q * NT] = P[co[q]];
}

// ------------------------------------------------------ density (plane pass)
// Density of a block of bands on the plane path: one workgroup per (z, band
// group, k). For each band of its group with a nonzero prefactor w (the
// filling times the k weight over the cell volume) it scatters the band's
// active columns at this z (T1 after the inverse z-transform, z-major) into
// the zeroed plane, transforms them back along y (active x rows) and x
// (every y), and adds w |psi|^2 to its own points, kept in registers in
// FP64 across the bands; the sum over its bands is added atomically to the
// z-major density rhoT[z][x][y].
// CROSS: the spinor cross term instead. T1 and T1b hold the two spinor
// components of the same bands; both planes are transformed (two LDS
// planes) and w psi_1 conj(psi_0) is accumulated, its real part into
// rhoT[z][x][y] and its imaginary part into rhoT[nz][...] above it.
template <typename C2, typename R, int NT, int MAXE, bool CROSS>
__global__ void __launch_bounds__(NT)
    k_plane_density(const C2* __restrict__ T1, const C2* __restrict__ T1b,
                    const double* __restrict__ prefac, int64_t ldp,
                    const int* __restrict__ col_ix, const int* __restrict__ col_iy,
                    const int* __restrict__ x_list, const C2* __restrict__ twx_g,
                    const C2* __restrict__ twy_g, Radices rx, Radices ry, FastDiv fny,
                    FastDiv fnxa, int ncol, int nxa, int nb, int b0, int bpg,
                    double* __restrict__ rhoT) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int kAcc = MAXE;  // points per thread: nx * ny <= NT * MAXE on the plane path
  constexpr int kNp = CROSS ? 2 : 1;  // planes in LDS
  const int nx = rx.n, ny = ry.n, ldp_ = odd_ld(ny), nz = gridDim.x;
  const int plane = nx * ldp_;
  C2* P = reinterpret_cast<C2*>(smem_raw);
  C2* twx = P + kNp * plane;
  C2* twy = twx + nx;
  int* xl = reinterpret_cast<int*>(twy + ny);
  const int z = blockIdx.x, g = blockIdx.y, k = blockIdx.z;
  const int tid = threadIdx.x, nxy = nx * ny;
  for (int i = tid; i < nx; i += NT) twx[i] = twx_g[i];
  for (int i = tid; i < ny; i += NT) twy[i] = twy_g[i];
  for (int i = tid; i < nxa; i += NT) xl[i] = x_list[i];
  double acc[kAcc], acc_im[CROSS ? kAcc : 1];
#pragma unroll
  for (int q = 0; q < kAcc; ++q) acc[q] = 0.0;
  if constexpr (CROSS) {
#pragma unroll
    for (int q = 0; q < kAcc; ++q) acc_im[q] = 0.0;
  }
  const int bend = min(nb, (g + 1) * bpg);
  for (int ib = g * bpg; ib < bend; ++ib) {
    const double w = prefac[(int64_t)k * ldp + b0 + ib];
    if (w == 0.0) continue;  // empty band (uniform over the workgroup)
    const int64_t roff = (((int64_t)k * nb + ib) * nz + z) * ncol;
    R cvx[kNp][kPlaneColPer], cvy[kNp][kPlaneColPer];  // real, imaginary (no scratch)
    int co[kPlaneColPer];
#pragma unroll
    for (int q = 0; q < kPlaneColPer; ++q) {
      const int c = tid + q * NT;
      if (c < ncol) {
        const C2 t = T1[roff + c];
        cvx[0][q] = t.x;
        cvy[0][q] = t.y;
        if constexpr (CROSS) {
          const C2 u = T1b[roff + c];
          cvx[1][q] = u.x;
          cvy[1][q] = u.y;
        }
        co[q] = col_ix[c] * ldp_ + col_iy[c];
      }
    }
    zero_lds(P, kNp * plane);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kPlaneColPer; ++q)
      if (tid + q * NT < ncol) {
        P[co[q]] = C2{cvx[0][q], cvy[0][q]};
        if constexpr (CROSS) P[plane + co[q]] = C2{cvx[1][q], cvy[1][q]};
      }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kNp; ++j) {
      fft_ip<C2, +1, NT, MAXE, false, true>(P + j * plane, twy, ry, nxa, fnxa, 1, ldp_, xl);
      fft_ip<C2, +1, NT, MAXE, true, false>(P + j * plane, twx, rx, ny, fny, ldp_, 1, nullptr);
    }
#pragma unroll
    for (int q = 0; q < kAcc; ++q) {
      const int i = tid + q * NT;
      if (i < nxy) {
        const int x = fdiv(i, fny);
        const C2 v = P[x * ldp_ + (i - x * ny)];
        if constexpr (CROSS) {  // psi_1 conj(psi_0), psi_0 = v, psi_1 = u
          const C2 u = P[plane + x * ldp_ + (i - x * ny)];
          acc[q] += w * ((double)u.x * (double)v.x + (double)u.y * (double)v.y);
          acc_im[q] += w * ((double)u.y * (double)v.x - (double)u.x * (double)v.y);
        } else {
          acc[q] += w * ((double)v.x * (double)v.x + (double)v.y * (double)v.y);
        }
      }
    }
    __syncthreads();  // the planes are zeroed for the next band
  }
  double* rz = rhoT + (int64_t)z * nxy;
#pragma unroll
  for (int q = 0; q < kAcc; ++q) {
    const int i = tid + q * NT;
    if (i < nxy && acc[q] != 0.0) unsafeAtomicAdd(rz + i, acc[q]);
    if constexpr (CROSS)
      if (i < nxy && acc_im[q] != 0.0) unsafeAtomicAdd(rz + (int64_t)nz * nxy + i, acc_im[q]);
  }
}

// rho[x][y][z] (= or +=) rhoT[z][x][y] (32 x 32 tiles through LDS) and a
// fill kernel for the z-major accumulator.
template <bool kAccumulate>
__global__ void __launch_bounds__(256)
    k_density_unmajor(const double* __restrict__ rhoT, double* __restrict__ rho, int nxy, int nz) {
  __shared__ double tile[32][33];
  const int xy0 = blockIdx.x * 32, z0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int r = ty; r < 32; r += 8) {
    const int zz = z0 + r, xy = xy0 + tx;
    if (xy < nxy && zz < nz) tile[r][tx] = rhoT[(int64_t)zz * nxy + xy];
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int xy = xy0 + r, zz = z0 + tx;
    if (xy < nxy && zz < nz) {
      double* o = rho + (int64_t)xy * nz + zz;
      *o = kAccumulate ? *o + tile[tx][r] : tile[tx][r];
    }
  }
}

__global__ void __launch_bounds__(256) k_fill_zero(double* __restrict__ x, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] = 0.0;
}

// The order of x and y within a plane that a fixed plane kernel asks of its
// potential (plane_fixed_space_order): the rho order n' = B ((n ia) mod A)
// + (n ib) mod B for a prime-factor size n = A B (ia = B^-1 mod A,
// ib = A^-1 mod B); A = 0 for the natural order.
struct PlaneOrder {
  int n, A, B, ia, ib;
  __device__ int rho(int v) const { return B * ((v * ia) % A) + (v * ib) % B; }
};

PlaneOrder plane_order(bool fixed, int n) {
  PlaneOrder po{n, 0, 1, 0, 0};
  int A, B;
  if (fixed && plane_fixed_space_order(n, &A, &B)) {
    po.A = A;
    po.B = B;
    po.ia = inv_mod(B, A);
    po.ib = inv_mod(A, B);
  }
  return po;
}

// Vt[s][z][xy] = V[s][xy][z]: each potential box to z-major order (32 x 32
// tiles through LDS, both sides coalesced), x and y in the order po asks.
template <typename R>
__global__ void __launch_bounds__(256)
    k_potential_zmajor(const R* __restrict__ V, R* __restrict__ Vt, int nxy, int nz,
                       PlaneOrder po) {
  __shared__ R tile[32][33];
  const int s = blockIdx.z, xy0 = blockIdx.x * 32, z0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const R* Vs = V + (int64_t)s * nxy * nz;
  R* Vts = Vt + (int64_t)s * nxy * nz;
  for (int r = ty; r < 32; r += 8) {
    const int xy = xy0 + r, zz = z0 + tx;
    if (xy < nxy && zz < nz) tile[r][tx] = Vs[(int64_t)xy * nz + zz];
  }
  __syncthreads();
  const int xy = xy0 + tx;
  int dst = xy;  // (x, y) in the rho order of the plane kernel where it asks for it
  if (po.A > 0 && xy < nxy) {
    const int x = xy / po.n, y = xy - x * po.n;
    dst = po.rho(x) * po.n + po.rho(y);
  }
  for (int r = ty; r < 32; r += 8) {
    const int zz = z0 + r;
    if (xy < nxy && zz < nz) Vts[(int64_t)zz * nxy + dst] = tile[tx][r];
  }
}

// ---------------------------------------------------------------- pass E
template <typename C2, typename R, bool ZMAJOR>
__global__ void __launch_bounds__(kThreads)
    k_cols_to_sphere(const C2* __restrict__ T1, const C2* C, int64_t ldb, int64_t ldk,
                     C2* out, int64_t oldb, int64_t oldk, const R* __restrict__ ke,
                     int64_t ke_ldk, const int* __restrict__ t1idx,
                     const int* __restrict__ cstart, const int* __restrict__ nvec,
                     const C2* __restrict__ tw_g, Radices rz, FastDiv fm, int ncol, int nb,
                     int n_tot, int b0, R scale, int accumulate) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int nz = rz.n, M = fm.d, ld = odd_ld(M);
  C2* tw = reinterpret_cast<C2*>(smem_raw);
  C2* a = tw + nz;
  C2* b = a + nz * ld;
  const int col0 = blockIdx.x * M, ib = blockIdx.y, k = blockIdx.z;
  const int mcols = min(M, ncol - col0);
  load_twiddles(tw, tw_g, nz);
  const C2* slab = T1 + ((int64_t)k * nb + ib) * ncol * nz;
  if (ZMAJOR) {
    for (int i = threadIdx.x; i < M * nz; i += blockDim.x) {
      const int p = fdiv(i, fm);
      const int l = i - p * M;
      a[p * ld + l] = l < mcols ? slab[(int64_t)p * ncol + col0 + l] : C2{0, 0};
    }
  } else {
    const C2* src = slab + (int64_t)col0 * nz;
    for (int i = threadIdx.x; i < M * nz; i += blockDim.x) {
      const int l = i / nz;
      const int p = i - l * nz;
      a[p * ld + l] = l < mcols ? src[i] : C2{0, 0};
    }
  }
  __syncthreads();
  C2* res = fft_lds<C2, -1>(a, b, tw, rz, fm, ld);
  const int* cs = cstart + (int64_t)k * (ncol + 1);
  const int gs = cs[col0], ge = cs[col0 + mcols];
  const int64_t band = b0 + ib;
  const C2* Cb = C + k * ldk + band * ldb;
  C2* ob = out + k * oldk + band * oldb;
  const R* keb = ke ? ke + k * ke_ldk : nullptr;
  const int* ti = t1idx + (int64_t)k * n_tot;
  const int base = col0 * nz;
  for (int g = gs + threadIdx.x; g < ge; g += blockDim.x) {
    const int t = ti[g] - base;
    const int l = t / nz;
    C2 v = cscale(res[(t - l * nz) * ld + l], scale);
    if (keb) {
      const C2 c = Cb[g];
      const R e = keb[g];
      v.x += e * c.x;
      v.y += e * c.y;
    }
    if (accumulate) v = cadd(v, ob[g]);
    ob[g] = v;
  }
  if (blockIdx.x == 0) {
    for (int g = nvec[k] + threadIdx.x; g < n_tot; g += blockDim.x) ob[g] = C2{0, 0};
  }
}

// ------------------------------------------------------------ dense FFT
// Lines along the contiguous last axis: M lines of length n per workgroup.
template <typename C2, typename R, int S>
__global__ void __launch_bounds__(kThreads)
    k_fft_contig(C2* __restrict__ data, int64_t n_lines, const C2* __restrict__ tw_g,
                 Radices rad, FastDiv fm, R scale) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int n = rad.n, M = fm.d, ld = odd_ld(M);
  C2* tw = reinterpret_cast<C2*>(smem_raw);
  C2* a = tw + n;
  C2* b = a + n * ld;
  const int64_t line0 = (int64_t)blockIdx.x * M;
  const int m = (int)min<int64_t>(M, n_lines - line0);
  load_twiddles(tw, tw_g, n);
  C2* base = data + line0 * n;
  for (int i = threadIdx.x; i < M * n; i += blockDim.x) {
    const int l = i / n;
    const int p = i - l * n;
    a[p * ld + l] = l < m ? base[i] : C2{0, 0};
  }
  __syncthreads();
  C2* res = fft_lds<C2, S>(a, b, tw, rad, fm, ld);
  for (int i = threadIdx.x; i < m * n; i += blockDim.x) {
    const int l = i / n;
    const int p = i - l * n;
    base[i] = cscale(res[p * ld + l], scale);
  }
}

// Lines along a strided axis, tiled over the contiguous last axis (length
// n_inner). Line o starts at (o / cnt) * s_hi + (o % cnt) * s_lo.
template <typename C2, typename R, int S>
__global__ void __launch_bounds__(kThreads)
    k_fft_strided(C2* __restrict__ data, int64_t line_stride, int n_inner, int ntile, int cnt,
                  int64_t s_hi, int64_t s_lo, const C2* __restrict__ tw_g, Radices rad,
                  FastDiv fm, R scale) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int n = rad.n, M = fm.d, ld = odd_ld(M);
  C2* tw = reinterpret_cast<C2*>(smem_raw);
  C2* a = tw + n;
  C2* b = a + n * ld;
  const int64_t o = blockIdx.x / ntile;
  const int z0 = (int)(blockIdx.x - o * ntile) * M;
  const int mz = min(M, n_inner - z0);
  C2* base = data + (o / cnt) * s_hi + (o % cnt) * s_lo + z0;
  load_twiddles(tw, tw_g, n);
  for (int i = threadIdx.x; i < n * M; i += blockDim.x) {
    const int p = fdiv(i, fm);
    const int l = i - p * M;
    a[p * ld + l] = l < mz ? base[p * line_stride + l] : C2{0, 0};
  }
  __syncthreads();
  C2* res = fft_lds<C2, S>(a, b, tw, rad, fm, ld);
  for (int i = threadIdx.x; i < n * M; i += blockDim.x) {
    const int p = fdiv(i, fm);
    const int l = i - p * M;
    if (l < mz) base[p * line_stride + l] = cscale(res[p * ld + l], scale);
  }
}

// Twiddle tables e^(-2 pi i t/n) in double, rounded once to float for c64.
template <typename C2>
static std::vector<C2> make_twiddles(int n) {
  std::vector<C2> out(n);
  for (int t = 0; t < n; ++t) {
    const long double ang = -2.0L * 3.141592653589793238462643383279502884L * t / n;
    out[t].x = (decltype(out[t].x))std::cos(ang);
    out[t].y = (decltype(out[t].x))std::sin(ang);
  }
  return out;
}

struct TwCache {
  std::mutex mu;
  // (device, n) -> tables; small and kept for the process lifetime.
  std::vector<std::pair<std::pair<int, int>, std::pair<float2*, double2*>>> entries;
};
TwCache g_tw;

static int get_twiddles(int n, float2** t32, double2** t64) {
  int dev = 0;
  QMX_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_tw.mu);
  for (auto& e : g_tw.entries) {
    if (e.first.first == dev && e.first.second == n) {
      *t32 = e.second.first;
      *t64 = e.second.second;
      return QMX_OK;
    }
  }
  auto h32 = make_twiddles<float2>(n);
  auto h64 = make_twiddles<double2>(n);
  float2* d32 = nullptr;
  double2* d64 = nullptr;
  QMX_HIP(hipMalloc(&d32, sizeof(float2) * n));
  QMX_HIP(hipMalloc(&d64, sizeof(double2) * n));
  QMX_HIP(hipMemcpy(d32, h32.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
  QMX_HIP(hipMemcpy(d64, h64.data(), sizeof(double2) * n, hipMemcpyHostToDevice));
  g_tw.entries.push_back({{dev, n}, {d32, d64}});
  *t32 = d32;
  *t64 = d64;
  return QMX_OK;
}

template <typename C2>
const C2* tw_of(const qmx_plan* p, int axis);
template <>
const float2* tw_of<float2>(const qmx_plan* p, int axis) {
  return p->d_tw32[axis];
}
template <>
const double2* tw_of<double2>(const qmx_plan* p, int axis) {
  return p->d_tw64[axis];
}

// Plane path (fused B-C-D) when a plane fits in LDS and every FFT pass
// fits the per-thread register budget; plan option 1 forces the separate
// passes, 2 asks for the plane path where it fits.
static size_t plane_lds_bytes(const qmx_plan* p, int cs) {
  const int nx = p->shape[0], ny = p->shape[1];
  return (size_t)nx * odd_ld(ny) * cs + (size_t)(nx + ny) * cs + (size_t)p->nxa * sizeof(int);
}

// Threads of the plane kernel for this box: 256 for planes of up to
// kPlaneSmallPts points whose passes fit the register budget, else 512.
static int plane_threads(const qmx_plan* p) {
  const int nt = kPlaneThreadsSmall;
  if ((int64_t)p->shape[0] * p->shape[1] <= kPlaneSmallPts && p->ncol <= kPlaneColPer * nt &&
      fft_ip_fits(p->rad[1], p->nxa, nt, kPlaneMaxESmall) &&
      fft_ip_fits(p->rad[0], p->shape[1], nt, kPlaneMaxESmall))
    return nt;
  return kPlaneThreads;
}

static bool plane_fits(const qmx_plan* p, int cs) {
  return p->nxa > 0 && plane_lds_bytes(p, cs) <= (size_t)kLdsMax &&
         p->ncol <= kPlaneColPer * kPlaneThreads &&
         fft_ip_fits(p->rad[1], p->nxa, kPlaneThreads, kPlaneMaxE) &&
         fft_ip_fits(p->rad[0], p->shape[1], kPlaneThreads, kPlaneMaxE);
}

static bool use_plane_path(const qmx_plan* p, int cs) {
  return p->local_path != 1 && plane_fits(p, cs);
}

static int64_t plane_aware_band_bytes(const qmx_plan* p, int cs, bool plane) {
  const int64_t nz = p->shape[2], ny = p->shape[1];
  return plane ? (int64_t)p->ncol * nz * cs : ((int64_t)p->ncol * nz + (int64_t)p->nxa * ny * nz) * cs;
}

static int64_t vt_bytes(const qmx_plan* p, int rs) {
  return ((int64_t)p->n_vslots * p->N * rs + 255) / 256 * 256;
}

template <typename C2, typename R, int NT, int MAXE>
static int set_plane_lds(size_t bytes) {
  static size_t done = 0;
  if (bytes > done) {
    QMX_HIP(hipFuncSetAttribute((const void*)k_plane_potential<C2, R, NT, MAXE>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    done = bytes;
  }
  return QMX_OK;
}

template <typename R>
int apply_local_impl(const qmx_plan* p, int k0, int nk, const void* Cv, int64_t ldb, int64_t ldk,
                     void* outv, int64_t oldb, int64_t oldk, const void* Vv, const void* kev,
                     int64_t ke_ldk, int64_t nb, int accumulate, void* ws, int64_t ws_bytes,
                     hipStream_t st) {
  using C2 = typename Cplx<R>::T;
  const int idoff = sizeof(R) == 8 ? QMX_KERNEL_C128 : 0;  // profile ids per precision
  const int nx = p->shape[0], ny = p->shape[1], nz = p->shape[2];
  const int ncol = p->ncol, nxa = p->nxa;
  const int* t1idx = p->d_t1idx + (int64_t)k0 * p->n_tot;
  const int* cstart = p->d_cstart + (int64_t)k0 * (ncol + 1);
  const int* nvec = p->d_n + k0;
  const int* vslot = p->d_vslot + k0;
  const bool plane = use_plane_path(p, sizeof(C2));
  const int64_t per_band = plane_aware_band_bytes(p, sizeof(C2), plane);
  const int64_t fixed = plane ? vt_bytes(p, sizeof(R)) : 0;
  QMX_CHECK_ARG(ws_bytes > fixed, "qmx_apply_local: workspace smaller than the potential boxes");
  int64_t blk = (ws_bytes - fixed) / (per_band * nk);
  QMX_CHECK_ARG(blk >= 1, "qmx_apply_local: workspace smaller than one band of every k");
  blk = std::min<int64_t>(blk, nb);
  blk = std::min<int64_t>(blk, 65535 / nk);  // grid.z of passes B and D is nk * blk
  QMX_CHECK_ARG(blk >= 1, "qmx_apply_local: too many k-points for one launch");
  const C2* C = static_cast<const C2*>(Cv);
  C2* out = static_cast<C2*>(outv);
  const R* V = static_cast<const R*>(Vv);
  const R* ke = static_cast<const R*>(kev);
  const int cs = sizeof(C2);
  // Tile sizes under the LDS budget:
  const int Mz = max_lines(nz, cs, 0, 64);
  const int My = even_tile(nz, max_lines(ny, cs, 0, 64));
  // Pass C: z-tile of Mzc (power of two, <= 16) and nbg bands per group,
  // M = nbg * Mzc lines, a multiple of the 4 waves, under the LDS budget.
  int Mzc = 16, nbg = 1;
  while (Mzc > 1 && Mzc / 2 >= nz) Mzc /= 2;
  for (;;) {
    int best = 0;
    for (int g = 8; g >= 1; g /= 2) {
      const int M = g * Mzc;
      const size_t need = (size_t)(2 * odd_ld(M) * nx + nx) * cs + (size_t)nx * Mzc * sizeof(R);
      if (M % 4 == 0 && need <= (size_t)kLdsBudget) { best = g; break; }
    }
    if (best) { nbg = best; break; }
    if (Mzc <= 4) { nbg = 4 / Mzc > 0 ? 4 / Mzc : 1; break; }
    Mzc /= 2;
  }
  const int Mx = nbg * Mzc;
  int zshift = 0;
  while ((1 << zshift) < Mzc) ++zshift;
  const FastDiv fz = make_fastdiv(Mz), fy = make_fastdiv(My), fxw = make_fastdiv(Mx / 4),
                fxM = make_fastdiv(Mx);
  const size_t shA = (size_t)(2 * odd_ld(Mz) * nz + nz) * cs;
  const size_t shB = (size_t)(2 * odd_ld(My) * ny + ny) * cs;
  const size_t shC = (size_t)(2 * odd_ld(Mx) * nx + nx) * cs + (size_t)nx * Mzc * sizeof(R);
  const R scale = R(1) / R(p->N);
  // Sphere points per band summed over k (for the byte counts):
  const double n_sphere = (double)(p->h_nsum[k0 + nk] - p->h_nsum[k0]);
  const double bc = (double)cs, br = (double)sizeof(R);
  const double t1_elems = (double)nk * ncol * nz, t2_elems = (double)nk * nxa * ny * nz;
  hipEvent_t ev = nullptr, ev_apply = nullptr;
  prof_begin(st, &ev_apply);
  const size_t shP = plane_lds_bytes(p, sizeof(C2));
  int ntP = plane ? plane_threads(p) : 0;
  bool fixed_ok = p->local_path != 3;  // the specialised plane kernel where it exists
  if (plane) {
    // Tuned plane kernel: 0 specialised, 1 general on 256 threads, 2 on 512.
    const int64_t key[4] = {sizeof(R) == 4 ? QMX_C64 : QMX_C128, nx, ny, nz};
    int64_t ch[1];
    if (tune_lookup(QMX_TUNE_PLANE, key, 4, ch, 1)) {
      if (ch[0] == 2) {
        ntP = kPlaneThreads;
        fixed_ok = false;
      } else if (ch[0] == 1 && ntP == kPlaneThreadsSmall) {
        fixed_ok = false;
      }
    }
  }
  const bool plane_fixed =
      plane && fixed_ok && ntP == kPlaneThreadsSmall && plane_fixed_available(nx, ny, ncol);
  const R* Vt = nullptr;
  if (plane) {
    // z-major copy of every potential box at the start of the workspace
    R* vt = static_cast<R*>(ws);
    const int nxy = nx * ny;
    dim3 gT((nxy + 31) / 32, (nz + 31) / 32, p->n_vslots);
    prof_begin(st, &ev);
    hipLaunchKernelGGL((k_potential_zmajor<R>), gT, dim3(256), 0, st, V, vt, nxy, nz,
                       plane_order(plane_fixed, nx));
    QMX_LAUNCHED();
    prof_end(st, ev, idoff + QMX_KERNEL_VZMAJOR, 2.0 * p->n_vslots * (double)p->N * br);
    Vt = vt;
  }
  if (plane) {
    const int rc = ntP == kPlaneThreadsSmall
                       ? set_plane_lds<C2, R, kPlaneThreadsSmall, kPlaneMaxESmall>(shP)
                       : set_plane_lds<C2, R, kPlaneThreads, kPlaneMaxE>(shP);
    if (rc) return rc;
  }
  const FastDiv fny = make_fastdiv(ny), fnxa = make_fastdiv(nxa);
  // The column passes with the z length fixed at compile time, where they
  // exist (path 3 forces the general kernels everywhere).
  const bool cols_fixed = plane && p->col_interval && p->local_path != 3 &&
                          cols_fixed_available(nz, sizeof(R) == 8);
  ColFixedArgs fa{};
  if (cols_fixed) {
    fa.st = st;
    fa.C = C;
    fa.ldb = ldb;
    fa.ldk = ldk;
    fa.out = out;
    fa.oldb = oldb;
    fa.oldk = oldk;
    fa.ke = ke;
    fa.ke_ldk = ke_ldk;
    fa.cstart = cstart;
    fa.zrange = p->d_zrange + (int64_t)k0 * ncol;
    fa.nvec = nvec;
    fa.tw = tw_of<C2>(p, 2);
    fa.ncol = ncol;
    fa.n_tot = p->n_tot;
    fa.run_max = p->col_run_max;
    fa.scale = (double)scale;
    fa.accumulate = accumulate;
  }
  for (int64_t b0 = 0; b0 < nb; b0 += blk) {
    const int cur = (int)std::min<int64_t>(blk, nb - b0);
    C2* T1 = reinterpret_cast<C2*>(static_cast<char*>(ws) + fixed);
    C2* T2 = T1 + (int64_t)nk * cur * ncol * nz;
    dim3 gA((ncol + Mz - 1) / Mz, cur, nk);
    fa.grid = dim3((ncol + kColFixedThreads - 1) / kColFixedThreads, cur, nk);
    fa.T1 = T1;
    fa.nb = cur;
    fa.b0 = (int)b0;
    prof_begin(st, &ev);
    if (cols_fixed) {
      const int rc = launch_cols_fixed(sizeof(R) == 8, false, nz, fa);
      if (rc) return rc;
    } else if (plane)
      hipLaunchKernelGGL((k_sphere_to_cols<C2, R, true>), gA, dim3(kThreads), shA, st, C, ldb,
                         ldk, T1, t1idx, cstart, tw_of<C2>(p, 2), p->rad[2], fz, ncol, cur,
                         p->n_tot, (int)b0);
    else
      hipLaunchKernelGGL((k_sphere_to_cols<C2, R, false>), gA, dim3(kThreads), shA, st, C, ldb,
                         ldk, T1, t1idx, cstart, tw_of<C2>(p, 2), p->rad[2], fz, ncol, cur,
                         p->n_tot, (int)b0);
    QMX_LAUNCHED();
    prof_end(st, ev, idoff + QMX_KERNEL_PASS_A, cur * (n_sphere * bc + t1_elems * bc));
    if (plane) {
      dim3 gP(nz, cur, nk);
      prof_begin(st, &ev);
      // Composite radices where they fit the plane kernel's register budget.
      const int maxeP = ntP == kPlaneThreadsSmall ? kPlaneMaxESmall : kPlaneMaxE;
      const Radices rxc = combine_radices(p->rad[0], ny, ntP, maxeP);
      const Radices ryc = combine_radices(p->rad[1], nxa, ntP, maxeP);
      if (plane_fixed) {
        const int rc = launch_plane_fixed(sizeof(R) == 8, nx, gP, shP, st, T1, Vt, vslot,
                                          p->d_col_ix, p->d_col_iy, p->d_x_list, tw_of<C2>(p, 0),
                                          tw_of<C2>(p, 1), ncol, nxa, cur);
        if (rc) return rc;
      } else if (ntP == kPlaneThreadsSmall)
        hipLaunchKernelGGL((k_plane_potential<C2, R, kPlaneThreadsSmall, kPlaneMaxESmall>), gP,
                           dim3(kPlaneThreadsSmall), shP, st, T1, Vt, vslot, p->d_col_ix,
                           p->d_col_iy, p->d_x_list, tw_of<C2>(p, 0), tw_of<C2>(p, 1), rxc,
                           ryc, fny, fnxa, ncol, nxa, cur);
      else
        hipLaunchKernelGGL((k_plane_potential<C2, R, kPlaneThreads, kPlaneMaxE>), gP, dim3(kPlaneThreads),
                           shP, st, T1, Vt, vslot, p->d_col_ix, p->d_col_iy, p->d_x_list,
                           tw_of<C2>(p, 0), tw_of<C2>(p, 1), rxc, ryc, fny, fnxa, ncol, nxa,
                           cur);
      QMX_LAUNCHED();
      prof_end(st, ev, idoff + QMX_KERNEL_PLANE, cur * 2.0 * t1_elems * bc + (double)nk * p->N * br);
    } else {
    dim3 gB((nz + My - 1) / My, nxa, nk * cur);
    prof_begin(st, &ev);
    hipLaunchKernelGGL((k_cols_to_planes<C2, R>), gB, dim3(kThreads), shB, st, T1, T2,
                       p->d_pcol_start, p->d_col_iy, tw_of<C2>(p, 1), p->rad[1], fy, ncol,
                       nxa, nz);
    QMX_LAUNCHED();
    prof_end(st, ev, idoff + QMX_KERNEL_PASS_B, cur * (t1_elems + t2_elems) * bc);
    // Bands per chunk: enough workgroups to fill the chip, V reused within.
    const int64_t wg_base = (int64_t)((nz + Mzc - 1) / Mzc) * ny * nk;
    int n_chunks = (int)std::min<int64_t>((cur + nbg - 1) / nbg,
                                          std::max<int64_t>(1, (2048 + wg_base - 1) / wg_base));
    n_chunks = std::min<int>(n_chunks, 65535 / nk);
    int bpc = (cur + n_chunks - 1) / n_chunks;
    bpc = (bpc + nbg - 1) / nbg * nbg;
    n_chunks = (cur + bpc - 1) / bpc;
    dim3 gC((nz + Mzc - 1) / Mzc, ny, nk * n_chunks);
    prof_begin(st, &ev);
    hipLaunchKernelGGL((k_xpass_potential<C2, R>), gC, dim3(kThreads), shC, st, T2, V, vslot,
                       p->d_x_list, tw_of<C2>(p, 0), p->rad[0], fxw, fxM, zshift, nbg, ny, nz,
                       nxa, cur, bpc, n_chunks);
    QMX_LAUNCHED();
    prof_end(st, ev, idoff + QMX_KERNEL_PASS_C,
             cur * 2.0 * t2_elems * bc + (double)nk * n_chunks * (double)p->N * br);
    prof_begin(st, &ev);
    hipLaunchKernelGGL((k_planes_to_cols<C2, R>), gB, dim3(kThreads), shB, st, T2, T1,
                       p->d_pcol_start, p->d_col_iy, tw_of<C2>(p, 1), p->rad[1], fy, ncol,
                       nxa, nz);
    QMX_LAUNCHED();
    prof_end(st, ev, idoff + QMX_KERNEL_PASS_D, cur * (t1_elems + t2_elems) * bc);
    }
    prof_begin(st, &ev);
    if (cols_fixed) {
      const int rc = launch_cols_fixed(sizeof(R) == 8, true, nz, fa);
      if (rc) return rc;
    } else if (plane)
      hipLaunchKernelGGL((k_cols_to_sphere<C2, R, true>), gA, dim3(kThreads), shA, st, T1, C,
                         ldb, ldk, out, oldb, oldk, ke, ke_ldk, t1idx, cstart, nvec,
                         tw_of<C2>(p, 2), p->rad[2], fz, ncol, cur, p->n_tot, (int)b0, scale,
                         accumulate);
    else
      hipLaunchKernelGGL((k_cols_to_sphere<C2, R, false>), gA, dim3(kThreads), shA, st, T1, C,
                         ldb, ldk, out, oldb, oldk, ke, ke_ldk, t1idx, cstart, nvec,
                         tw_of<C2>(p, 2), p->rad[2], fz, ncol, cur, p->n_tot, (int)b0, scale,
                         accumulate);
    QMX_LAUNCHED();
    prof_end(st, ev, idoff + QMX_KERNEL_PASS_E,
             cur * (t1_elems * bc + n_sphere * bc * (1 + (ke ? 1 : 0) + (accumulate ? 1 : 0)) +
                    (ke ? n_sphere * br : 0.0)));
  }
  // The whole apply under the dense-transform convention: per band a sphere
  // read and write plus one box read and write for each of the two 3D
  // transforms, and the potential read once per k-point.
  prof_end(st, ev_apply, idoff + QMX_KERNEL_APPLY,
           (double)nb * (2.0 * n_sphere * bc + 4.0 * nk * (double)p->N * bc) +
               (double)nk * p->N * br);
  return QMX_OK;
}

template <typename C2, typename R, int NT, int MAXE, bool CROSS>
static int set_density_lds(size_t bytes) {
  static size_t done = 0;
  if (bytes > done) {
    QMX_HIP(hipFuncSetAttribute((const void*)k_plane_density<C2, R, NT, MAXE, CROSS>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    done = bytes;
  }
  return QMX_OK;
}

// Workspace of the density: the z-major accumulator (N doubles, 2 N for the
// spinor cross term), then T1 of a block of bands (of both components for
// the cross term).
static int64_t density_fixed_bytes(const qmx_plan* p, bool cross) {
  return ((int64_t)p->N * 8 * (cross ? 2 : 1) + 255) / 256 * 256;
}

template <typename C2, typename R, int NT, int MAXE, bool CROSS>
static void launch_density(dim3 grid, size_t sh, hipStream_t st, const C2* T1, const C2* T1b,
                           const double* prefac, int64_t ldp, const qmx_plan* p, Radices rxc,
                           Radices ryc, FastDiv fny, FastDiv fnxa, int cur, int b0, int bpg,
                           double* rhoT) {
  hipLaunchKernelGGL((k_plane_density<C2, R, NT, MAXE, CROSS>), grid, dim3(NT), sh, st, T1, T1b,
                     prefac, ldp, p->d_col_ix, p->d_col_iy, p->d_x_list, tw_of<C2>(p, 0),
                     tw_of<C2>(p, 1), rxc, ryc, fny, fnxa, p->ncol, p->nxa, cur, b0, bpg, rhoT);
}

// Whether the density of this box runs on the plane pass (x-y planes, two
// for the spinor cross term, in LDS and within the accumulator registers);
// otherwise on the separate passes.
static bool density_plane(const qmx_plan* p, int cs, bool cross) {
  if (!use_plane_path(p, cs)) return false;
  const size_t shP = plane_lds_bytes(p, cs) + (cross ? (size_t)p->shape[0] * odd_ld(p->shape[1]) * cs : 0);
  const int ntP = plane_threads(p);
  const int maxeP = ntP == kPlaneThreadsSmall ? kPlaneMaxESmall : kPlaneMaxE;
  return shP <= (size_t)kLdsMax && (int64_t)p->shape[0] * p->shape[1] <= (int64_t)ntP * maxeP;
}

// Workspace per band and k of the density (T1, plus T2 on the separate passes).
static int64_t density_band_bytes(const qmx_plan* p, int cs, bool cross) {
  return plane_aware_band_bytes(p, cs, density_plane(p, cs, cross)) * (cross ? 2 : 1);
}

template <typename C2, typename R, bool CROSS>
static int set_xdensity_lds(size_t bytes) {
  static size_t done = 0;
  if (bytes > done) {
    QMX_HIP(hipFuncSetAttribute((const void*)k_xpass_density<C2, R, CROSS>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    done = bytes;
  }
  return QMX_OK;
}

// The density on the separate passes: passes A and B of the local apply
// (z- then y-transforms into T2) for each component, then k_xpass_density;
// rho (= or +=) the x-major accumulator.
template <typename R>
int collect_density_separate(const qmx_plan* p, int k0, int nk, const void* Cv, const void* Cv1,
                             int64_t ldb, int64_t ldk, int64_t nb, const double* prefac,
                             int64_t ldp, double* rho, int accumulate, void* ws, int64_t ws_bytes,
                             hipStream_t st, const char* name) {
  using C2 = typename Cplx<R>::T;
  const bool cross = Cv1 != nullptr;
  const int n_comp = cross ? 2 : 1;
  const int nx = p->shape[0], ny = p->shape[1], nz = p->shape[2];
  const int ncol = p->ncol, nxa = p->nxa, cs = sizeof(C2);
  const int* t1idx = p->d_t1idx + (int64_t)k0 * p->n_tot;
  const int* cstart = p->d_cstart + (int64_t)k0 * (ncol + 1);
  const int64_t fixed = density_fixed_bytes(p, cross);
  const int64_t per_band = density_band_bytes(p, cs, cross);
  if (ws_bytes <= fixed || (ws_bytes - fixed) / (per_band * nk) < 1) {
    set_error(std::string(name) + ": workspace smaller than the density and one band of every k");
    return QMX_ERR_ARG;
  }
  int64_t blk = (ws_bytes - fixed) / (per_band * nk);
  blk = std::min<int64_t>(blk, nb);
  blk = std::min<int64_t>(blk, 65535 / nk);
  // Pass C tiles: z-tile Mzc (power of two, <= 16), nbg bands per group,
  // M = nbg * Mzc lines (a multiple of the 4 waves) under the LDS budget.
  auto need_c = [&](int M, int Mzc) {
    return (size_t)(2 * n_comp * odd_ld(M) * nx + nx) * cs + (size_t)n_comp * nx * Mzc * 8;
  };
  int Mzc = 16, nbg = 0;
  while (Mzc > 1 && Mzc / 2 >= nz) Mzc /= 2;
  for (size_t budget : {(size_t)kLdsBudget, (size_t)kLdsMax}) {
    for (int mz = Mzc; mz >= 1 && !nbg; mz /= 2)
      for (int g = 8; g >= 1; g /= 2)
        if ((g * mz) % 4 == 0 && need_c(g * mz, mz) <= budget) {
          Mzc = mz;
          nbg = g;
          break;
        }
    if (nbg) break;
  }
  if (!nbg) {
    set_error(std::string(name) + ": x lines of this box exceed LDS");
    return QMX_ERR_ARG;
  }
  const int Mx = nbg * Mzc;
  int zshift = 0;
  while ((1 << zshift) < Mzc) ++zshift;
  const int Mz = max_lines(nz, cs, 0, 64);
  const int My = even_tile(nz, max_lines(ny, cs, 0, 64));
  const FastDiv fz = make_fastdiv(Mz), fy = make_fastdiv(My), fxw = make_fastdiv(Mx / 4),
                fxM = make_fastdiv(Mx);
  const size_t shA = (size_t)(2 * odd_ld(Mz) * nz + nz) * cs;
  const size_t shB = (size_t)(2 * odd_ld(My) * ny + ny) * cs;
  const size_t shC = need_c(Mx, Mzc);
  int rc = cross ? set_xdensity_lds<C2, R, true>(shC) : set_xdensity_lds<C2, R, false>(shC);
  if (rc) return rc;
  const C2* Cs[2] = {static_cast<const C2*>(Cv), static_cast<const C2*>(Cv1)};
  double* rhoX = static_cast<double*>(ws);
  const int64_t n_acc = (int64_t)p->N * n_comp;
  hipLaunchKernelGGL(k_fill_zero, dim3(std::min<int64_t>((n_acc + 255) / 256, 4096)), dim3(256), 0,
                     st, rhoX, n_acc);
  QMX_LAUNCHED();
  const double* pf = prefac + k0 * ldp;
  for (int64_t b0 = 0; b0 < nb; b0 += blk) {
    const int cur = (int)std::min<int64_t>(blk, nb - b0);
    const int64_t t1n = (int64_t)cur * nk * ncol * nz, t2n = (int64_t)cur * nk * nxa * ny * nz;
    C2* T1 = reinterpret_cast<C2*>(static_cast<char*>(ws) + fixed);
    C2* T2[2] = {T1 + n_comp * t1n, T1 + n_comp * t1n + t2n};
    for (int c = 0; c < n_comp; ++c) {
      C2* T1c = T1 + c * t1n;
      dim3 gA((ncol + Mz - 1) / Mz, cur, nk);
      hipLaunchKernelGGL((k_sphere_to_cols<C2, R, false>), gA, dim3(kThreads), shA, st, Cs[c], ldb,
                         ldk, T1c, t1idx, cstart, tw_of<C2>(p, 2), p->rad[2], fz, ncol, cur,
                         p->n_tot, (int)b0);
      QMX_LAUNCHED();
      dim3 gB((nz + My - 1) / My, nxa, nk * cur);
      hipLaunchKernelGGL((k_cols_to_planes<C2, R>), gB, dim3(kThreads), shB, st, T1c, T2[c],
                         p->d_pcol_start, p->d_col_iy, tw_of<C2>(p, 1), p->rad[1], fy, ncol, nxa,
                         nz);
      QMX_LAUNCHED();
    }
    const int64_t wg_base = (int64_t)((nz + Mzc - 1) / Mzc) * ny * nk;
    int n_chunks = (int)std::min<int64_t>((cur + nbg - 1) / nbg,
                                          std::max<int64_t>(1, (2048 + wg_base - 1) / wg_base));
    n_chunks = std::min<int>(n_chunks, 65535 / nk);
    int bpc = (cur + n_chunks - 1) / n_chunks;
    bpc = (bpc + nbg - 1) / nbg * nbg;
    n_chunks = (cur + bpc - 1) / bpc;
    dim3 gC((nz + Mzc - 1) / Mzc, ny, nk * n_chunks);
    if (cross)
      hipLaunchKernelGGL((k_xpass_density<C2, R, true>), gC, dim3(kThreads), shC, st, T2[0], T2[1],
                         pf, ldp, p->d_x_list, tw_of<C2>(p, 0), p->rad[0], fxw, fxM, zshift, nbg,
                         ny, nz, nxa, cur, (int)b0, bpc, n_chunks, rhoX);
    else
      hipLaunchKernelGGL((k_xpass_density<C2, R, false>), gC, dim3(kThreads), shC, st, T2[0],
                         nullptr, pf, ldp, p->d_x_list, tw_of<C2>(p, 0), p->rad[0], fxw, fxM,
                         zshift, nbg, ny, nz, nxa, cur, (int)b0, bpc, n_chunks, rhoX);
    QMX_LAUNCHED();
  }
  const unsigned gD = (unsigned)std::min<int64_t>((n_acc + 255) / 256, 8192);
  if (accumulate)
    hipLaunchKernelGGL(k_density_add<true>, dim3(gD), dim3(256), 0, st, rhoX, rho, n_acc);
  else
    hipLaunchKernelGGL(k_density_add<false>, dim3(gD), dim3(256), 0, st, rhoX, rho, n_acc);
  QMX_LAUNCHED();
  return QMX_OK;
}

// Cv1 == nullptr: rho (N doubles) = or += sum_b w_b |psi_b|^2. Otherwise the
// spinor cross term: rho (2 N doubles, real then imaginary part) = or +=
// sum_b w_b psi1_b conj(psi0_b), psi0 in Cv, psi1 in Cv1 (same strides).
template <typename R>
int collect_density_impl(const qmx_plan* p, int k0, int nk, const void* Cv, const void* Cv1,
                         int64_t ldb, int64_t ldk, int64_t nb, const double* prefac, int64_t ldp,
                         double* rho, int accumulate, void* ws, int64_t ws_bytes, hipStream_t st) {
  using C2 = typename Cplx<R>::T;
  const bool cross = Cv1 != nullptr;
  const int n_comp = cross ? 2 : 1;
  const int nx = p->shape[0], ny = p->shape[1], nz = p->shape[2];
  const int ncol = p->ncol, nxa = p->nxa, cs = sizeof(C2);
  const char* name = cross ? "qmx_collect_density_cross" : "qmx_collect_density";
  if (!density_plane(p, cs, cross))
    return collect_density_separate<R>(p, k0, nk, Cv, Cv1, ldb, ldk, nb, prefac, ldp, rho,
                                       accumulate, ws, ws_bytes, st, name);
  const size_t shP = plane_lds_bytes(p, cs) + (cross ? (size_t)nx * odd_ld(ny) * cs : 0);
  const int* t1idx = p->d_t1idx + (int64_t)k0 * p->n_tot;
  const int* cstart = p->d_cstart + (int64_t)k0 * (ncol + 1);
  const int64_t fixed = density_fixed_bytes(p, cross);
  const int64_t per_band = (int64_t)ncol * nz * cs * n_comp;
  if (ws_bytes <= fixed || (ws_bytes - fixed) / (per_band * nk) < 1) {
    set_error(std::string(name) + ": workspace smaller than the density and one band of every k");
    return QMX_ERR_ARG;
  }
  int64_t blk = (ws_bytes - fixed) / (per_band * nk);
  blk = std::min<int64_t>(blk, nb);
  blk = std::min<int64_t>(blk, 65535 / nk);
  const C2* Cs[2] = {static_cast<const C2*>(Cv), static_cast<const C2*>(Cv1)};
  double* rhoT = static_cast<double*>(ws);
  const int64_t n_acc = (int64_t)p->N * n_comp;
  hipLaunchKernelGGL(k_fill_zero, dim3(std::min<int64_t>((n_acc + 255) / 256, 4096)), dim3(256), 0,
                     st, rhoT, n_acc);
  QMX_LAUNCHED();
  const int Mz = max_lines(nz, cs, 0, 64);
  const FastDiv fz = make_fastdiv(Mz);
  const size_t shA = (size_t)(2 * odd_ld(Mz) * nz + nz) * cs;
  const int ntP = plane_threads(p);
  const bool small = ntP == kPlaneThreadsSmall;
  const int maxeP = small ? kPlaneMaxESmall : kPlaneMaxE;
  int rc = cross ? (small ? set_density_lds<C2, R, kPlaneThreadsSmall, kPlaneMaxESmall, true>(shP)
                          : set_density_lds<C2, R, kPlaneThreads, kPlaneMaxE, true>(shP))
                 : (small ? set_density_lds<C2, R, kPlaneThreadsSmall, kPlaneMaxESmall, false>(shP)
                          : set_density_lds<C2, R, kPlaneThreads, kPlaneMaxE, false>(shP));
  if (rc) return rc;
  const Radices rxc = combine_radices(p->rad[0], ny, ntP, maxeP);
  const Radices ryc = combine_radices(p->rad[1], nxa, ntP, maxeP);
  const FastDiv fny = make_fastdiv(ny), fnxa = make_fastdiv(nxa);
  for (int64_t b0 = 0; b0 < nb; b0 += blk) {
    const int cur = (int)std::min<int64_t>(blk, nb - b0);
    C2* T1[2];
    T1[0] = reinterpret_cast<C2*>(static_cast<char*>(ws) + fixed);
    T1[1] = T1[0] + (int64_t)cur * nk * ncol * nz;
    dim3 gA((ncol + Mz - 1) / Mz, cur, nk);
    for (int c = 0; c < n_comp; ++c) {
      hipLaunchKernelGGL((k_sphere_to_cols<C2, R, true>), gA, dim3(kThreads), shA, st, Cs[c], ldb,
                         ldk, T1[c], t1idx, cstart, tw_of<C2>(p, 2), p->rad[2], fz, ncol, cur,
                         p->n_tot, (int)b0);
      QMX_LAUNCHED();
    }
    // Band groups: enough workgroups to fill the chip, each accumulating
    // several bands in registers.
    const int64_t wg_base = (int64_t)nz * nk;
    int groups = (int)std::min<int64_t>(cur, std::max<int64_t>(1, (2048 + wg_base - 1) / wg_base));
    const int bpg = (cur + groups - 1) / groups;
    groups = (cur + bpg - 1) / bpg;
    dim3 gP(nz, groups, nk);
    const double* pf = prefac + k0 * ldp;
    if (cross) {
      if (small)
        launch_density<C2, R, kPlaneThreadsSmall, kPlaneMaxESmall, true>(
            gP, shP, st, T1[0], T1[1], pf, ldp, p, rxc, ryc, fny, fnxa, cur, (int)b0, bpg, rhoT);
      else
        launch_density<C2, R, kPlaneThreads, kPlaneMaxE, true>(
            gP, shP, st, T1[0], T1[1], pf, ldp, p, rxc, ryc, fny, fnxa, cur, (int)b0, bpg, rhoT);
    } else {
      if (small)
        launch_density<C2, R, kPlaneThreadsSmall, kPlaneMaxESmall, false>(
            gP, shP, st, T1[0], nullptr, pf, ldp, p, rxc, ryc, fny, fnxa, cur, (int)b0, bpg, rhoT);
      else
        launch_density<C2, R, kPlaneThreads, kPlaneMaxE, false>(
            gP, shP, st, T1[0], nullptr, pf, ldp, p, rxc, ryc, fny, fnxa, cur, (int)b0, bpg, rhoT);
    }
    QMX_LAUNCHED();
  }
  const int nxy = nx * ny;
  dim3 gT((nxy + 31) / 32, (nz + 31) / 32);
  for (int c = 0; c < n_comp; ++c) {
    const double* src = rhoT + (int64_t)c * p->N;
    double* dst = rho + (int64_t)c * p->N;
    if (accumulate)
      hipLaunchKernelGGL(k_density_unmajor<true>, gT, dim3(256), 0, st, src, dst, nxy, nz);
    else
      hipLaunchKernelGGL(k_density_unmajor<false>, gT, dim3(256), 0, st, src, dst, nxy, nz);
    QMX_LAUNCHED();
  }
  return QMX_OK;
}

template <typename R>
int fft3d_impl(const int32_t* shape, void* datav, int64_t batch, int direction, double scale,
               hipStream_t st) {
  using C2 = typename Cplx<R>::T;
  C2* data = static_cast<C2*>(datav);
  const int n0 = shape[0], n1 = shape[1], n2 = shape[2];
  const int cs = sizeof(C2);
  for (int axis = 2; axis >= 0; --axis) {
    const int n = shape[axis];
    Radices rad;
    QMX_CHECK_ARG(factorize(n, &rad), "qmx_fft3d: box size must be a product of 2, 3 and 5");
    float2* t32;
    double2* t64;
    int rc = get_twiddles(n, &t32, &t64);
    if (rc) return rc;
    const C2* tw = reinterpret_cast<const C2*>(sizeof(R) == 4 ? (const void*)t32 : (const void*)t64);
    const R sc = axis == 0 ? R(scale) : R(1);
    if (axis == 2) {
      const int M = max_lines(n, cs, 0, 64);
      const FastDiv fm = make_fastdiv(M);
      const int64_t n_lines = batch * n0 * n1;
      const size_t sh = (size_t)(2 * odd_ld(M) * n + n) * cs;
      const int64_t nblk = (n_lines + M - 1) / M;
      QMX_CHECK_ARG(nblk < (1ll << 31), "qmx_fft3d: batch too large");
      if (direction < 0)
        hipLaunchKernelGGL((k_fft_contig<C2, R, -1>), dim3((unsigned)nblk), dim3(kThreads), sh, st,
                           data, n_lines, tw, rad, fm, sc);
      else
        hipLaunchKernelGGL((k_fft_contig<C2, R, +1>), dim3((unsigned)nblk), dim3(kThreads), sh, st,
                           data, n_lines, tw, rad, fm, sc);
    } else {
      const int M = even_tile(n2, max_lines(n, cs, 0, 64));
      const FastDiv fm = make_fastdiv(M);
      const int ntile = (n2 + M - 1) / M;
      int64_t n_outer, s_hi, s_lo, line_stride;
      int cnt;
      if (axis == 1) {
        n_outer = batch * n0;
        cnt = 1;
        s_hi = (int64_t)n1 * n2;
        s_lo = 0;
        line_stride = n2;
      } else {
        n_outer = batch * n1;
        cnt = n1;
        s_hi = (int64_t)n0 * n1 * n2;
        s_lo = n2;
        line_stride = (int64_t)n1 * n2;
      }
      const size_t sh = (size_t)(2 * odd_ld(M) * n + n) * cs;
      const int64_t nblk = n_outer * ntile;
      QMX_CHECK_ARG(nblk < (1ll << 31), "qmx_fft3d: batch too large");
      if (direction < 0)
        hipLaunchKernelGGL((k_fft_strided<C2, R, -1>), dim3((unsigned)nblk), dim3(kThreads), sh, st,
                           data, line_stride, n2, ntile, cnt, s_hi, s_lo, tw, rad, fm, sc);
      else
        hipLaunchKernelGGL((k_fft_strided<C2, R, +1>), dim3((unsigned)nblk), dim3(kThreads), sh, st,
                           data, line_stride, n2, ntile, cnt, s_hi, s_lo, tw, rad, fm, sc);
    }
    QMX_LAUNCHED();
  }
  return QMX_OK;
}

// Table blocks of destroyed plans (a device block, its pinned staging
// copy and the event recorded after the copy between them) are kept for
// reuse by later plans: hipFree and hipHostMalloc wait for or serialise
// with the device, and a server creates and destroys a plan whenever a
// bucket's membership changes. A block is reused only for a plan of at
// most its size and at least half of it; at most kPoolMax bytes of free
// blocks are kept (beyond that the largest ones are released).
struct PoolBlock {
  void* d;
  void* h;
  size_t bytes;
  hipEvent_t uploaded;
  hipEvent_t last_use;  // after the last kernel of the previous owner that read d
};
static std::mutex g_pool_mu;
static std::vector<PoolBlock> g_pool;
static size_t g_pool_bytes = 0;
constexpr size_t kPoolMax = size_t(1) << 30;
constexpr size_t kPoolGrain = size_t(1) << 16;

static int pool_acquire(size_t need, PoolBlock* out) {
  {
    std::lock_guard<std::mutex> lock(g_pool_mu);
    int best = -1;
    for (int i = 0; i < (int)g_pool.size(); ++i) {
      const size_t b = g_pool[i].bytes;
      if (b >= need && b <= 2 * need + kPoolGrain && (best < 0 || b < g_pool[best].bytes)) best = i;
    }
    if (best >= 0) {
      *out = g_pool[best];
      g_pool_bytes -= out->bytes;
      g_pool.erase(g_pool.begin() + best);
    } else {
      out->d = nullptr;
    }
  }
  if (out->d) {
    // The staging copy is rewritten: its previous copy to the device must
    // have finished (long ago in practice).
    QMX_HIP(hipEventSynchronize(out->uploaded));
    return QMX_OK;
  }
  const size_t bytes = (need + kPoolGrain - 1) / kPoolGrain * kPoolGrain;
  *out = PoolBlock{nullptr, nullptr, bytes, nullptr, nullptr};
  QMX_HIP(hipHostMalloc(&out->h, bytes, hipHostMallocDefault));
  QMX_HIP(hipMalloc(&out->d, bytes));
  QMX_HIP(hipEventCreateWithFlags(&out->uploaded, hipEventDisableTiming));
  QMX_HIP(hipEventCreateWithFlags(&out->last_use, hipEventDisableTiming));
  return QMX_OK;
}

static void pool_release(const PoolBlock& blk) {
  std::vector<PoolBlock> drop;
  {
    std::lock_guard<std::mutex> lock(g_pool_mu);
    g_pool.push_back(blk);
    g_pool_bytes += blk.bytes;
    while (g_pool_bytes > kPoolMax && !g_pool.empty()) {
      auto it = std::max_element(g_pool.begin(), g_pool.end(),
                                 [](const PoolBlock& a, const PoolBlock& b) { return a.bytes < b.bytes; });
      drop.push_back(*it);
      g_pool_bytes -= it->bytes;
      g_pool.erase(it);
    }
  }
  for (auto& b : drop) {
    (void)hipFree(b.d);
    (void)hipHostFree(b.h);
    (void)hipEventDestroy(b.uploaded);
    (void)hipEventDestroy(b.last_use);
  }
}

// The integer tables of a plan go to the device as one block: staged in
// pinned host memory and copied asynchronously on the null stream, so
// creating a plan does not wait for the work already queued on the device.
// A reused block may still be read by kernels of its previous plan on any
// stream (non-blocking streams do not order with the null stream): the copy
// first waits for the event recorded after that plan's last call. Every
// entry point that reads a plan waits for `uploaded` on its own stream
// (plan_enter).
struct TableSpan {
  const int* src;  // copied into the block; null: written by the fill callback
  size_t n;        // ints
  int** dst;       // set to the table's device address
};

template <typename Fill>
static int upload_tables(qmx_plan* p, const std::vector<TableSpan>& spans, Fill&& fill) {
  std::vector<size_t> off(spans.size());
  size_t total = 0;
  for (size_t i = 0; i < spans.size(); ++i) {
    off[i] = total;
    total += (spans[i].n * sizeof(int) + 255) / 256 * 256;
    *spans[i].dst = nullptr;
  }
  if (total == 0) return QMX_OK;
  PoolBlock blk;
  const int rc = pool_acquire(total, &blk);
  if (rc) return rc;
  p->d_block = blk.d;
  p->h_stage = blk.h;
  p->block_bytes = blk.bytes;
  p->uploaded = blk.uploaded;
  p->last_use = blk.last_use;
  auto* h = static_cast<char*>(p->h_stage);
  auto* d = static_cast<char*>(p->d_block);
  std::vector<int*> host(spans.size());
  for (size_t i = 0; i < spans.size(); ++i) {
    host[i] = reinterpret_cast<int*>(h + off[i]);
    if (spans[i].n == 0) continue;
    if (spans[i].src) std::memcpy(host[i], spans[i].src, spans[i].n * sizeof(int));
    *spans[i].dst = reinterpret_cast<int*>(d + off[i]);
  }
  fill(host);
  QMX_HIP(hipStreamWaitEvent(nullptr, p->last_use, 0));
  QMX_HIP(hipMemcpyAsync(d, h, total, hipMemcpyHostToDevice, nullptr));
  QMX_HIP(hipEventRecord(p->uploaded, nullptr));
  return QMX_OK;
}

// Every entry point that reads a plan's device tables brackets its launches
// on stream st with these: before them, st waits for the tables' upload and,
// when the previous call ran on another stream, for that call's work (so a
// plan moved between streams stays ordered); after them, last_use marks the
// point a reused block's next upload must wait for.
static int plan_enter(const qmx_plan* cp, hipStream_t st) {
  auto* p = const_cast<qmx_plan*>(cp);
  if (p->entered && p->last_stream == st) return QMX_OK;
  QMX_HIP(hipStreamWaitEvent(st, p->uploaded, 0));
  if (p->entered) QMX_HIP(hipStreamWaitEvent(st, p->last_use, 0));
  p->entered = true;
  p->last_stream = st;
  return QMX_OK;
}

static int plan_leave(const qmx_plan* p, hipStream_t st, int rc) {
  if (p->last_use) {
    const hipError_t e = hipEventRecord(p->last_use, st);
    if (e != hipSuccess && rc == QMX_OK) {
      set_error(std::string("hipEventRecord: ") + hipGetErrorString(e));
      return QMX_ERR_HIP;
    }
  }
  return rc;
}

static void free_plan(qmx_plan* p) {
  if (!p) return;
  if (p->d_block)
    pool_release(PoolBlock{p->d_block, p->h_stage, p->block_bytes, p->uploaded, p->last_use});
  delete[] p->h_nsum;
  delete p;
}

}  // namespace

extern "C" {

const char* qmx_last_error(void) { return g_last_error.c_str(); }
const char* qmx_version(void) { return "qmx 0.1 gfx950"; }

}  // extern "C"

namespace {

// Host-side index tables of a basis (no device calls), built in two phases:
// the column structure (scan_columns), then the per-k tables into arrays the
// caller provides (fill_k_tables), so that a plan writes them straight into
// its pinned staging block. Both phases split the k-points over threads
// when a geometry has enough sphere points to pay for them.
struct GeomColumns {
  std::vector<int> x_list, pcol_start, col_iy, col_ix;
  std::vector<int> col_of;  // [n0 * n1] column index of each box column, -1 if empty
};

struct KTables {
  int* cstart;        // [nk][ncol + 1] first sphere point of each column
  int* t1idx;         // [nk][n_tot] col * n2 + z of each sphere point, 0 past n[k]
  int* zrange;        // [nk][ncol] packed z interval of each column (below)
  bool col_interval;  // every column of every k is one such interval
  int col_run_max;    // most points of kColFixedThreads consecutive columns of one k
};

constexpr int64_t kPointsPerThread = int64_t(1) << 18;
constexpr int kMaxTableThreads = 16;

int table_threads(int nk, int64_t points) {
  const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
  const int64_t t = std::min<int64_t>(std::min(hw, kMaxTableThreads), nk);
  return (int)std::max<int64_t>(1, std::min<int64_t>(t, points / kPointsPerThread + 1));
}

// f(slot, k_begin, k_end) over T consecutive ranges of [0, nk), range 0 on
// the calling thread (and any range whose thread cannot be started).
template <typename F>
void run_k_ranges(int T, int nk, F&& f) {
  auto lo = [&](int t) { return (int)((int64_t)nk * t / T); };
  std::vector<std::thread> th;
  th.reserve(T > 1 ? T - 1 : 0);
  for (int t = 1; t < T; ++t) {
    try {
      th.emplace_back([&f, t, a = lo(t), b = lo(t + 1)] { f(t, a, b); });
    } catch (...) {
      f(t, lo(t), lo(t + 1));
    }
  }
  f(0, 0, lo(1));
  for (auto& x : th) x.join();
}

const int64_t* fft_row(const qmx_geom* geom, int k) {
  return geom->fft_rows ? geom->fft_rows[k] : geom->fft_index + (int64_t)k * geom->n_tot;
}

int scan_columns(const qmx_geom* geom, GeomColumns* c) {
  QMX_CHECK_ARG(geom, "qmx_geom_tables: null geometry");
  const int n0 = geom->shape[0], n1 = geom->shape[1], n2 = geom->shape[2];
  const int nk = geom->nk, n_tot = geom->n_tot;
  QMX_CHECK_ARG(n0 > 0 && n1 > 0 && n2 > 0 && n0 <= 512 && n1 <= 512 && n2 <= 512,
                "qmx_geom_tables: box sizes must be in 1..512");
  QMX_CHECK_ARG(nk >= 1 && nk < 65536 && n_tot >= 1, "qmx_geom_tables: bad nk / n_tot");
  QMX_CHECK_ARG(geom->n && (geom->fft_index || geom->fft_rows),
                "qmx_geom_tables: null n / fft_index");
  if (geom->fft_rows)
    for (int k = 0; k < nk; ++k)
      QMX_CHECK_ARG(geom->fft_rows[k] || geom->n[k] == 0, "qmx_geom_tables: null fft_rows[k]");
  const int64_t N = (int64_t)n0 * n1 * n2;
  QMX_CHECK_ARG(N < (1ll << 31), "qmx_geom_tables: box too large");
  for (int a = 0; a < 3; ++a) {
    Radices rad;
    QMX_CHECK_ARG(factorize(geom->shape[a], &rad),
                  "qmx_geom_tables: box sizes must be products of 2, 3 and 5");
  }
  for (int k = 0; k < nk; ++k)
    QMX_CHECK_ARG(geom->n[k] >= 0 && geom->n[k] <= n_tot,
                  "qmx_geom_tables: n[k] outside 0..n_tot");
  // Columns touched by any k (a bitmap per thread, merged below). Sphere
  // points come in increasing box order, so one column's points are
  // consecutive: the column index is recomputed only where a point leaves
  // the current column.
  const int T = table_threads(nk, (int64_t)nk * n_tot);
  const size_t ncols_box = (size_t)n0 * n1;
  std::vector<std::vector<char>> seen(T);
  std::vector<char> bad(T, 0);
  run_k_ranges(T, nk, [&](int t, int k0, int k1) {
    std::vector<char>& s = seen[t];
    s.assign(ncols_box, 0);
    for (int k = k0; k < k1; ++k) {
      const int64_t* fi = fft_row(geom, k);
      int64_t prev = -1, col_end = -1;
      for (int g = 0; g < geom->n[k]; ++g) {
        const int64_t idx = fi[g];
        if (idx <= prev || idx >= N) {
          bad[t] = 1;
          return;
        }
        prev = idx;
        if (idx >= col_end) {
          const int64_t col = idx / n2;
          col_end = (col + 1) * n2;
          s[col] = 1;
        }
      }
    }
  });
  for (int t = 0; t < T; ++t)
    QMX_CHECK_ARG(!bad[t], "qmx_geom_tables: fft_index must be increasing box indices within each k");
  for (int t = 1; t < T; ++t)
    for (size_t i = 0; i < ncols_box; ++i) seen[0][i] |= seen[t][i];
  const std::vector<char>& col_seen = seen[0];
  c->col_of.assign(ncols_box, -1);
  c->x_list.clear();
  c->pcol_start.clear();
  c->col_iy.clear();
  c->col_ix.clear();
  for (int x = 0; x < n0; ++x) {
    const char* row = col_seen.data() + (size_t)x * n1;
    if (std::find(row, row + n1, 1) == row + n1) continue;
    c->x_list.push_back(x);
    c->pcol_start.push_back((int)c->col_iy.size());
    for (int y = 0; y < n1; ++y) {
      if (!row[y]) continue;
      c->col_of[(size_t)x * n1 + y] = (int)c->col_iy.size();
      c->col_iy.push_back(y);
      c->col_ix.push_back(x);
    }
  }
  c->pcol_start.push_back((int)c->col_iy.size());
  return QMX_OK;
}

// The per-k tables. The z interval of each column's points (zrange, for the
// fixed column passes): a sphere meets a line of the box in one interval of
// signed z, stored in increasing z as z = lo..hi (lo >= 0) or, when it
// contains z = 0 and negative z, as 0..hi then n + lo..n - 1 (lo < 0).
// Packed as (lo & 0xffff) | hi << 16; an empty column is (1, 0).
// col_interval turns false if any column's points are not one such interval.
void fill_k_tables(const qmx_geom* geom, const GeomColumns& c, KTables* out) {
  const int n2 = geom->shape[2], nk = geom->nk, n_tot = geom->n_tot;
  const int ncol = (int)c.col_iy.size();
  const int T = table_threads(nk, (int64_t)nk * n_tot);
  std::vector<char> interval(T, 1);
  std::vector<int> run_max(T, 0);
  run_k_ranges(T, nk, [&](int t, int k0, int k1) {
    std::vector<int> counts(ncol);
    bool ok = true;
    int rmax = 0;
    for (int k = k0; k < k1; ++k) {
      const int64_t* fi = fft_row(geom, k);
      int* cs = out->cstart + (size_t)k * (ncol + 1);
      int* t1 = out->t1idx + (size_t)k * n_tot;
      int* zr = out->zrange + (size_t)k * ncol;
      std::fill(counts.begin(), counts.end(), 0);
      int64_t col_begin = 0, col_end = -1;
      int cc = 0;
      const int nkk = geom->n[k];
      for (int g = 0; g < nkk; ++g) {
        const int64_t idx = fi[g];
        if (idx >= col_end) {
          const int64_t col = idx / n2;
          col_begin = col * n2;
          col_end = col_begin + n2;
          cc = c.col_of[col];
        }
        counts[cc]++;
        t1[g] = cc * n2 + (int)(idx - col_begin);
      }
      std::fill(t1 + nkk, t1 + n_tot, 0);
      cs[0] = 0;
      for (int j = 0; j < ncol; ++j) cs[j + 1] = cs[j] + counts[j];
      for (int j = 0; j < ncol; ++j) {
        const int a = cs[j], b = cs[j + 1], cnt = b - a, base = j * n2;
        int lo = 1, hi = 0;
        if (cnt > 0) {
          const int zf = t1[a] - base, zl = t1[b - 1] - base;
          if (zl - zf + 1 == cnt) {
            lo = zf;
            hi = zl;
          } else {
            int h = 0;  // points 0..h, then l..n2 - 1
            while (h + 1 < cnt && t1[a + h + 1] - base == h + 1) ++h;
            const int l = h + 1 < cnt ? t1[a + h + 1] - base : n2;
            if (zf == 0 && zl == n2 - 1 && cnt - (h + 1) == n2 - l) {
              lo = l - n2;
              hi = h;
            } else {
              ok = false;
            }
          }
        }
        zr[j] = (lo & 0xFFFF) | (hi << 16);
      }
      for (int c0 = 0; c0 < ncol; c0 += kColFixedThreads)
        rmax = std::max(rmax, cs[std::min(ncol, c0 + kColFixedThreads)] - cs[c0]);
    }
    interval[t] = ok;
    run_max[t] = rmax;
  });
  out->col_interval = true;
  out->col_run_max = 0;
  for (int t = 0; t < T; ++t) {
    out->col_interval = out->col_interval && interval[t];
    out->col_run_max = std::max(out->col_run_max, run_max[t]);
  }
}

template <typename T>
void copy_out(T* dst, const std::vector<T>& src) {
  if (dst && !src.empty()) std::memcpy(dst, src.data(), sizeof(T) * src.size());
}

}  // namespace

extern "C" {

int qmx_geom_tables(const qmx_geom* geom, int64_t* sizes, int32_t* x_list,
                    int32_t* pcol_start, int32_t* col_iy, int32_t* cstart, int32_t* t1idx) {
  QMX_CHECK_ARG(sizes, "qmx_geom_tables: null sizes");
  GeomColumns c;
  int rc = scan_columns(geom, &c);
  if (rc) return rc;
  sizes[0] = (int64_t)c.x_list.size();
  sizes[1] = (int64_t)c.col_iy.size();
  copy_out(x_list, c.x_list);
  copy_out(pcol_start, c.pcol_start);
  copy_out(col_iy, c.col_iy);
  if (cstart || t1idx) {
    const size_t nk = geom->nk, ncol = c.col_iy.size();
    std::vector<int> cs(nk * (ncol + 1)), t1(nk * geom->n_tot), zr(nk * ncol);
    KTables kt{cs.data(), t1.data(), zr.data(), true, 0};
    fill_k_tables(geom, c, &kt);
    copy_out(cstart, cs);
    copy_out(t1idx, t1);
  }
  return QMX_OK;
}

int qmx_plan_create(const qmx_geom* geom, qmx_plan** plan_out) {
  QMX_CHECK_ARG(geom && plan_out, "qmx_plan_create: null argument");
  *plan_out = nullptr;
  GeomColumns t;
  int rc = scan_columns(geom, &t);
  if (rc) return rc;
  auto* p = new qmx_plan();
  std::memset(p, 0, sizeof(qmx_plan));
  for (int a = 0; a < 3; ++a) {
    p->shape[a] = geom->shape[a];
    factorize(p->shape[a], &p->rad[a]);
  }
  p->nk = geom->nk;
  p->n_tot = geom->n_tot;
  p->N = (int64_t)p->shape[0] * p->shape[1] * p->shape[2];
  p->nxa = (int)t.x_list.size();
  p->h_nsum = new int64_t[geom->nk + 1];
  p->h_nsum[0] = 0;
  for (int k = 0; k < geom->nk; ++k) p->h_nsum[k + 1] = p->h_nsum[k] + geom->n[k];
  p->n_sum = p->h_nsum[geom->nk];
  p->ncol = (int)t.col_iy.size();
  std::vector<int> nvec(geom->n, geom->n + geom->nk);
  std::vector<int> vslot(geom->nk, 0);
  p->n_vslots = 1;
  if (geom->v_slot) {
    for (int k = 0; k < geom->nk; ++k) {
      if (geom->v_slot[k] < 0) {
        free_plan(p);
        set_error("qmx_plan_create: negative v_slot");
        return QMX_ERR_ARG;
      }
      vslot[k] = geom->v_slot[k];
      p->n_vslots = std::max(p->n_vslots, vslot[k] + 1);
    }
  }
  const size_t nk = geom->nk, ncol = t.col_iy.size();
  KTables kt{nullptr, nullptr, nullptr, true, 0};
  rc = upload_tables(p,
                     {{t.x_list.data(), t.x_list.size(), &p->d_x_list},
                      {t.pcol_start.data(), t.pcol_start.size(), &p->d_pcol_start},
                      {t.col_iy.data(), ncol, &p->d_col_iy},
                      {nullptr, nk * (ncol + 1), &p->d_cstart},
                      {nullptr, nk * geom->n_tot, &p->d_t1idx},
                      {nvec.data(), nk, &p->d_n},
                      {vslot.data(), nk, &p->d_vslot},
                      {t.col_ix.data(), ncol, &p->d_col_ix},
                      {nullptr, nk * ncol, &p->d_zrange}},
                     [&](const std::vector<int*>& h) {
                       kt.cstart = h[3];
                       kt.t1idx = h[4];
                       kt.zrange = h[8];
                       fill_k_tables(geom, t, &kt);
                     });
  if (rc) {
    free_plan(p);
    return rc;
  }
  p->col_interval = kt.col_interval;
  p->col_run_max = kt.col_run_max;
  hipError_t e = hipGetDevice(&p->device);
  if (e != hipSuccess) {
    free_plan(p);
    set_error(std::string("hipGetDevice: ") + hipGetErrorString(e));
    return QMX_ERR_HIP;
  }
  for (int a = 0; a < 3; ++a) {
    if ((rc = get_twiddles(p->shape[a], &p->d_tw32[a], &p->d_tw64[a]))) {
      free_plan(p);
      return rc;
    }
  }
  *plan_out = p;
  return QMX_OK;
}

int qmx_plan_destroy(qmx_plan* plan) {
  free_plan(plan);
  return QMX_OK;
}

int64_t qmx_plan_workspace_bytes(const qmx_plan* p, int dtype, int64_t n_bands) {
  if (!p || (dtype != QMX_C64 && dtype != QMX_C128) || n_bands < 0) return -1;
  const int cs = dtype == QMX_C64 ? 8 : 16;
  const bool plane = use_plane_path(p, cs);
  const int64_t fixed = plane ? vt_bytes(p, cs / 2) : 0;
  return fixed + plane_aware_band_bytes(p, cs, plane) * p->nk * n_bands;
}

int qmx_plan_set_local_path(qmx_plan* p, int path) {
  QMX_CHECK_ARG(p && path >= 0 && path <= 3, "qmx_plan_set_local_path: path must be 0..3");
  p->local_path = path;
  return QMX_OK;
}

int qmx_plan_local_path(const qmx_plan* p, int dtype) {
  if (!p || (dtype != QMX_C64 && dtype != QMX_C128)) return -1;
  return use_plane_path(p, dtype == QMX_C64 ? 8 : 16) ? 2 : 1;
}

int qmx_plan_info(const qmx_plan* p, int64_t* info) {
  QMX_CHECK_ARG(p && info, "qmx_plan_info: null argument");
  info[0] = p->nxa;
  info[1] = p->ncol;
  return QMX_OK;
}

int qmx_apply_local_k(const qmx_plan* plan, int64_t k0, int64_t nk, int dtype, const void* C,
                      int64_t ldb, int64_t ldk, void* out, int64_t out_ldb, int64_t out_ldk,
                      const void* V, const void* ke, int64_t ke_ldk, int64_t n_bands,
                      int accumulate, void* workspace, int64_t workspace_bytes, void* stream) {
  QMX_CHECK_ARG(plan && C && out && V, "qmx_apply_local: null argument");
  QMX_CHECK_ARG(n_bands >= 0, "qmx_apply_local: negative n_bands");
  QMX_CHECK_ARG(k0 >= 0 && nk >= 0 && k0 + nk <= plan->nk, "qmx_apply_local: k range outside plan");
  if (n_bands == 0 || nk == 0) return QMX_OK;
  QMX_CHECK_ARG(workspace && workspace_bytes > 0, "qmx_apply_local: workspace required");
  QMX_CHECK_ARG(ldb >= plan->n_tot && out_ldb >= plan->n_tot, "qmx_apply_local: ldb < n_tot");
  QMX_CHECK_ARG(nk == 1 || (ldk > 0 && out_ldk > 0), "qmx_apply_local: bad k stride");
  QMX_CHECK_ARG(!ke || nk == 1 || ke_ldk >= plan->n_tot, "qmx_apply_local: bad ke stride");
  QMX_CHECK_ARG(dtype == QMX_C64 || dtype == QMX_C128,
                "qmx_apply_local: dtype must be QMX_C64 or QMX_C128");
  hipStream_t st = static_cast<hipStream_t>(stream);
  int rc = plan_enter(plan, st);
  if (rc) return rc;
  if (dtype == QMX_C64)
    rc = apply_local_impl<float>(plan, (int)k0, (int)nk, C, ldb, ldk, out, out_ldb, out_ldk, V,
                                 ke, ke_ldk, n_bands, accumulate, workspace, workspace_bytes, st);
  else
    rc = apply_local_impl<double>(plan, (int)k0, (int)nk, C, ldb, ldk, out, out_ldb, out_ldk, V,
                                  ke, ke_ldk, n_bands, accumulate, workspace, workspace_bytes, st);
  return plan_leave(plan, st, rc);
}

int64_t qmx_density_workspace_bytes(const qmx_plan* plan, int dtype, int64_t n_bands) {
  if (!plan || (dtype != QMX_C64 && dtype != QMX_C128) || n_bands < 0) return -1;
  const int cs = dtype == QMX_C64 ? 8 : 16;
  return density_fixed_bytes(plan, false) + density_band_bytes(plan, cs, false) * plan->nk * n_bands;
}

int64_t qmx_density_cross_workspace_bytes(const qmx_plan* plan, int dtype, int64_t n_bands) {
  if (!plan || (dtype != QMX_C64 && dtype != QMX_C128) || n_bands < 0) return -1;
  const int cs = dtype == QMX_C64 ? 8 : 16;
  return density_fixed_bytes(plan, true) + density_band_bytes(plan, cs, true) * plan->nk * n_bands;
}

static int collect_density_entry(const char* name, const qmx_plan* plan, int64_t k0, int64_t nk,
                                 int dtype, const void* C0, const void* C1, int64_t ldb,
                                 int64_t ldk, int64_t n_bands, const double* prefac, int64_t ldp,
                                 double* rho, int accumulate, void* workspace,
                                 int64_t workspace_bytes, void* stream) {
  const std::string n(name);
  if (!plan || !C0 || !prefac || !rho) {
    set_error(n + ": null argument");
    return QMX_ERR_ARG;
  }
  if (n_bands < 0 || ldp < n_bands) {
    set_error(n + ": bad n_bands or ldp");
    return QMX_ERR_ARG;
  }
  if (k0 < 0 || nk < 0 || k0 + nk > plan->nk) {
    set_error(n + ": k range outside plan");
    return QMX_ERR_ARG;
  }
  if (!workspace || workspace_bytes <= 0) {
    set_error(n + ": workspace required");
    return QMX_ERR_ARG;
  }
  if (ldb < plan->n_tot || (nk > 1 && ldk <= 0)) {
    set_error(n + ": bad strides");
    return QMX_ERR_ARG;
  }
  if (dtype != QMX_C64 && dtype != QMX_C128) {
    set_error(n + ": dtype must be QMX_C64 or QMX_C128");
    return QMX_ERR_ARG;
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (n_bands == 0 || nk == 0) {
    if (accumulate) return QMX_OK;
    const int64_t len = (int64_t)plan->N * (C1 ? 2 : 1);
    hipLaunchKernelGGL(k_fill_zero, dim3(std::min<int64_t>((len + 255) / 256, 4096)), dim3(256),
                       0, st, rho, len);
    QMX_LAUNCHED();
    return QMX_OK;
  }
  int rc = plan_enter(plan, st);
  if (rc) return rc;
  if (dtype == QMX_C64)
    rc = collect_density_impl<float>(plan, (int)k0, (int)nk, C0, C1, ldb, ldk, n_bands, prefac,
                                     ldp, rho, accumulate, workspace, workspace_bytes, st);
  else
    rc = collect_density_impl<double>(plan, (int)k0, (int)nk, C0, C1, ldb, ldk, n_bands, prefac,
                                      ldp, rho, accumulate, workspace, workspace_bytes, st);
  return plan_leave(plan, st, rc);
}

int qmx_collect_density(const qmx_plan* plan, int64_t k0, int64_t nk, int dtype, const void* C,
                        int64_t ldb, int64_t ldk, int64_t n_bands, const double* prefac,
                        int64_t ldp, double* rho, int accumulate, void* workspace,
                        int64_t workspace_bytes, void* stream) {
  return collect_density_entry("qmx_collect_density", plan, k0, nk, dtype, C, nullptr, ldb, ldk,
                               n_bands, prefac, ldp, rho, accumulate, workspace, workspace_bytes,
                               stream);
}

int qmx_collect_density_cross(const qmx_plan* plan, int64_t k0, int64_t nk, int dtype,
                              const void* C0, const void* C1, int64_t ldb, int64_t ldk,
                              int64_t n_bands, const double* prefac, int64_t ldp, double* rho,
                              int accumulate, void* workspace, int64_t workspace_bytes,
                              void* stream) {
  if (!C1) {
    set_error("qmx_collect_density_cross: null argument");
    return QMX_ERR_ARG;
  }
  return collect_density_entry("qmx_collect_density_cross", plan, k0, nk, dtype, C0, C1, ldb,
                               ldk, n_bands, prefac, ldp, rho, accumulate, workspace,
                               workspace_bytes, stream);
}

int qmx_apply_local(const qmx_plan* plan, int dtype, const void* C, int64_t ldb, int64_t ldk,
                    void* out, int64_t out_ldb, int64_t out_ldk, const void* V, const void* ke,
                    int64_t ke_ldk, int64_t n_bands, int accumulate, void* workspace,
                    int64_t workspace_bytes, void* stream) {
  QMX_CHECK_ARG(plan, "qmx_apply_local: null argument");
  return qmx_apply_local_k(plan, 0, plan->nk, dtype, C, ldb, ldk, out, out_ldb, out_ldk, V, ke,
                           ke_ldk, n_bands, accumulate, workspace, workspace_bytes, stream);
}

int qmx_fft3d(int dtype, const int32_t* shape, void* data, int64_t batch, int direction,
              double scale, void* stream) {
  QMX_CHECK_ARG(shape && data, "qmx_fft3d: null argument");
  QMX_CHECK_ARG(direction == 1 || direction == -1, "qmx_fft3d: direction must be +-1");
  QMX_CHECK_ARG(batch >= 0, "qmx_fft3d: negative batch");
  for (int a = 0; a < 3; ++a)
    QMX_CHECK_ARG(shape[a] >= 1 && shape[a] <= 512, "qmx_fft3d: box sizes must be in 1..512");
  if (batch == 0) return QMX_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == QMX_C64) return fft3d_impl<float>(shape, data, batch, direction, scale, st);
  if (dtype == QMX_C128) return fft3d_impl<double>(shape, data, batch, direction, scale, st);
  set_error("qmx_fft3d: dtype must be QMX_C64 or QMX_C128");
  return QMX_ERR_ARG;
}


int qmx_profile_enable(int on) {
  std::lock_guard<std::mutex> lock(g_prof_mu);
  g_prof_on = on != 0;
  return QMX_OK;
}

int qmx_profile_read_flops(double* ms, double* bytes, double* flops, int64_t* count, int n_ids) {
  QMX_CHECK_ARG(ms && bytes && count && n_ids > 0, "qmx_profile_read: null argument");
  for (int i = 0; i < n_ids; ++i) {
    ms[i] = 0;
    bytes[i] = 0;
    if (flops) flops[i] = 0;
    count[i] = 0;
  }
  std::vector<ProfRec> recs;
  {
    std::lock_guard<std::mutex> lock(g_prof_mu);
    recs.swap(g_prof);
  }
  int rc = QMX_OK;
  for (auto& r : recs) {
    float t = 0;
    if (rc == QMX_OK) {
      hipError_t e = hipEventSynchronize(r.b);
      if (e == hipSuccess) e = hipEventElapsedTime(&t, r.a, r.b);
      if (e != hipSuccess) {
        set_error(std::string("qmx_profile_read: ") + hipGetErrorString(e));
        rc = QMX_ERR_HIP;
      }
    }
    if (r.id >= 0 && r.id < n_ids) {
      ms[r.id] += t;
      bytes[r.id] += r.bytes;
      if (flops) flops[r.id] += r.flops;
      count[r.id] += 1;
    }
    prof_release(r.a);
    prof_release(r.b);
  }
  return rc;
}

int qmx_profile_read(double* ms, double* bytes, int64_t* count, int n_ids) {
  return qmx_profile_read_flops(ms, bytes, nullptr, count, n_ids);
}

}  // extern "C"
