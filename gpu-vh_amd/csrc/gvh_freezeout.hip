// gvh_freezeout.hip -- freeze-out hypersurface of a running event on the GPU.
//
// The reference keeps FOFREQ+1 time slices of e and of ten hydro variables
// in host memory, refreshed after every step from a full device-to-host copy,
// and every FOFREQ steps runs cornelius serially over every cube of the
// window (HydroPlugin.cpp:183-393, FreezeOut.cpp:49-126).  Here the history
// ring lives in HBM next to the state and the window is searched by kernels:
//
//  1. fo_record_kernel: one slice of (u^x, u^y, u^eta, e, pi^xx, pi^xy,
//     pi^xeta, pi^yy, pi^yeta, Pi) from the ghosted state into the ring
//     (interior cells, x fastest).  The ring's slot 0 <- slot FOFREQ copy of
//     the reference is a swap of the logical -> physical slot map.
//  2. fo_rows_kernel (count, then list): one thread per (it, ix, iy) row of
//     the window walks eta, so neighbouring threads read neighbouring cells;
//     counts per row, an exclusive scan and a second walk give the surface
//     cubes in the reference's (it, ix, iy, ieta) order.
//  3. fo_cornelius_kernel: one wave per surface cube, elements via the LDS
//     restatement of cornelius (gvh_cornelius.h).
//  4. fo_emit_kernel: one thread per (element, variable) writes the rows
//     (position, normal, multilinearly interpolated hydro variables, T and P
//     of the interpolated e), offsets from a scan of element counts.
//
// Row format (iS3D, HydroPlugin.cpp:317-386): tau x y eta, t dsigma_tau
// t dsigma_x t dsigma_y t dsigma_eta, u^x u^y u^eta, e T P, pi^xx pi^xy
// pi^xeta pi^yy pi^yeta, Pi.  With flags 0 the reference's quirks are kept:
// the normal is scaled by the current time t, windows after the first place
// cell tau one step late, and after an odd number of steps the recorded u
// and pi are the previous step's (the reference copies them to the host
// through the buffers allocated first, which the pointer swaps leave one step
// behind on odd steps; e is updated in place and is always current).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <mutex>
#include <vector>

#include "../../include/gpuvh.h"
#include "gvh_cornelius.h"
#include "gvh_device.h"
#include "gvh_internal.h"

#pragma clang fp contract(off)

using namespace gvh;

namespace {

constexpr int NVAR = 10;      // ring variables (HydroPlugin.cpp:192-194)
constexpr int VAR_E = 3;
constexpr int NCOL = GVH_FO_COLUMNS;
constexpr int ELEM = 8;       // centroid[4], normal[4] per element
constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 8; // per thread
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

struct RingSlots {
  int phys[GVH_FO_MAX_FREQ + 1];
};

}  // namespace

namespace gvh {

struct FoState {
  int dim = 0, fofreq = 0, flags = 0;
  int halves = 1;  // ring copies of fofreq + 1 slots (2: a batch alternates them by window)
  double e_fo = 0;
  long long nint = 0;
  double *ring = nullptr;
  std::vector<int> slot;  // logical -> physical
  // window buffers, grown on demand
  unsigned *rowcnt = nullptr, *rowoff = nullptr;
  size_t row_cap = 0;           // rowcnt and rowoff
  unsigned *cubes = nullptr;     // cubes, ecnt, eoff, elem: cube_cap
  int *ecnt = nullptr;
  unsigned *eoff = nullptr;
  double *elem = nullptr;
  size_t cube_cap = 0;
  double *rows = nullptr;
  size_t rows_cap = 0;
  unsigned *scan_tmp = nullptr;
  size_t scan_cap = 0;
  unsigned *total = nullptr;  // device scalars: [0] scan total, [1] overflow
  long long ncubes = 0, nrows = 0;
  // the window search in progress (fo_find_count .. fo_find_emit)
  unsigned htot[2] = {0, 0};  // host copies of total[0], total[1]
  long long nrow_pending = 0;
  int n_pending = 0;
};

double *fo_ring(FoState *f) { return f ? f->ring : nullptr; }

void fo_free(FoState *f) {
  if (!f) return;
  for (void *p : {(void *)f->ring, (void *)f->rowcnt, (void *)f->rowoff, (void *)f->cubes,
                  (void *)f->ecnt, (void *)f->eoff, (void *)f->elem, (void *)f->rows,
                  (void *)f->scan_tmp, (void *)f->total})
    if (p) (void)hipFree(p);
  delete f;
}

}  // namespace gvh

// ------------------------------------------------------------------ scan
// Exclusive prefix sum of unsigned counts: tiles of SCAN_TILE, tile sums
// scanned recursively, tiles re-scanned with their offsets.
namespace {

__device__ unsigned block_exclusive_scan(unsigned v, unsigned *lds, unsigned &total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned x = v;
  for (int d = 1; d < 64; d <<= 1) {
    const unsigned y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) lds[wave] = x;
  __syncthreads();
  unsigned base = 0, tot = 0;
  for (int w = 0; w < SCAN_BLOCK / 64; ++w) {
    const unsigned s = lds[w];
    if (w < wave) base += s;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return base + x - v;
}

__global__ __launch_bounds__(SCAN_BLOCK) void scan_tiles_kernel(const unsigned *in, size_t n,
                                                                unsigned *tile_sums) {
  __shared__ unsigned lds[SCAN_BLOCK / 64];
  const size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
  unsigned s = 0;
  for (int k = 0; k < SCAN_ITEMS; ++k)
    if (base + k < n) s += in[base + k];
  unsigned tot;
  block_exclusive_scan(s, lds, tot);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SCAN_BLOCK) void scan_apply_kernel(const unsigned *in, size_t n,
                                                                const unsigned *tile_off,
                                                                unsigned *out) {
  __shared__ unsigned lds[SCAN_BLOCK / 64];
  const size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
  unsigned v[SCAN_ITEMS], s = 0;
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    v[k] = base + k < n ? in[base + k] : 0u;
    s += v[k];
  }
  unsigned tot;
  unsigned run = block_exclusive_scan(s, lds, tot) + (tile_off ? tile_off[blockIdx.x] : 0u);
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
}

__global__ void scan_total_kernel(const unsigned *in, const unsigned *out, size_t n,
                                  unsigned *total) {
  total[0] = n ? out[n - 1] + in[n - 1] : 0u;
}

// scratch words needed by exclusive_scan for n items
size_t scan_scratch(size_t n) {
  size_t s = 0;
  while (n > SCAN_TILE) {
    n = (n + SCAN_TILE - 1) / SCAN_TILE;
    s += 2 * n;
  }
  return s + 1;
}

hipError_t exclusive_scan(const unsigned *in, unsigned *out, size_t n, unsigned *scratch,
                          hipStream_t st) {
  if (n == 0) return hipSuccess;
  const size_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (tiles == 1) {
    scan_apply_kernel<<<1, SCAN_BLOCK, 0, st>>>(in, n, nullptr, out);
    return hipGetLastError();
  }
  unsigned *sums = scratch, *offs = scratch + tiles;
  scan_tiles_kernel<<<(unsigned)tiles, SCAN_BLOCK, 0, st>>>(in, n, sums);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = exclusive_scan(sums, offs, tiles, scratch + 2 * tiles, st);
  if (e != hipSuccess) return e;
  scan_apply_kernel<<<(unsigned)tiles, SCAN_BLOCK, 0, st>>>(in, n, offs, out);
  return hipGetLastError();
}

hipError_t scan_with_total(const unsigned *in, unsigned *out, size_t n, unsigned *scratch,
                           unsigned *total, hipStream_t st) {
  hipError_t e = exclusive_scan(in, out, n, scratch, st);
  if (e != hipSuccess) return e;
  scan_total_kernel<<<1, 1, 0, st>>>(in, out, n, total);
  return hipGetLastError();
}

// ---------------------------------------------------------------- record
// pi and Pi are read from qP: 11 blocks of G values of the storage type PS
template <typename PS>
__global__ __launch_bounds__(256) void fo_record_kernel(const double *__restrict__ e,
                                                        const double *__restrict__ u,
                                                        const PS *__restrict__ qP, long long G,
                                                        int nx, int ny, int nz, int ncx, int ncy,
                                                        int bulk, double *__restrict__ slot,
                                                        long long nint) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nint) return;
  const int i = (int)(c % nx);
  const long long r = c / nx;
  const int j = (int)(r % ny), k = (int)(r / ny);
  const long long s = (i + 2) + (long long)ncx * ((j + 2) + (long long)ncy * (k + 2));
  auto pi = [&](int f) { return (double)qP[(f - GVH_PITT) * G + s]; };
  slot[0 * nint + c] = u[1 * G + s];
  slot[1 * nint + c] = u[2 * G + s];
  slot[2 * nint + c] = u[3 * G + s];
  slot[3 * nint + c] = e[s];
  slot[4 * nint + c] = pi(GVH_PIXX);
  slot[5 * nint + c] = pi(GVH_PIXY);
  slot[6 * nint + c] = pi(GVH_PIXN);
  slot[7 * nint + c] = pi(GVH_PIYY);
  slot[8 * nint + c] = pi(GVH_PIYN);
  slot[9 * nint + c] = bulk ? pi(GVH_PI) : 0.0;
}

// ------------------------------------------------------- window geometry
struct Window {
  const double *ring;
  RingSlots sl;
  long long nint;
  int nx, ny, nz, dim, dimz;
  int start, nit;       // it = start .. start + nit - 1
  double e_fo;
  __device__ const double *e_slot(int it) const {
    return ring + ((long long)sl.phys[it] * NVAR + VAR_E) * nint;
  }
  __device__ long long cell(int ix, int iy, int iz) const {
    return ix + (long long)nx * (iy + (long long)ny * iz);
  }
};

// LIST == false: rowcnt[r] = surface cubes of row r; LIST == true: their ids
// at rowoff[r] onwards.  Row r = (itl (nx-1) + ix) (ny-1) + iy, cube id
// r dimz + iz (the reference's loop order, HydroPlugin.cpp:283-299).
template <bool LIST>
__global__ __launch_bounds__(256) void fo_rows_kernel(Window W, unsigned *__restrict__ rowcnt,
                                                      const unsigned *__restrict__ rowoff,
                                                      unsigned *__restrict__ cubes) {
  const long long nrow = (long long)(W.nx - 1) * (W.ny - 1) * W.nit;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nrow) return;
  const int ix = (int)(t % (W.nx - 1));
  const long long rest = t / (W.nx - 1);
  const int iy = (int)(rest % (W.ny - 1));
  const int itl = (int)(rest / (W.ny - 1));
  const long long r = ((long long)itl * (W.nx - 1) + ix) * (W.ny - 1) + iy;
  const double *e0 = W.e_slot(W.start + itl), *e1 = W.e_slot(W.start + itl + 1);
  const int ncorner = 1 << W.dim;
  unsigned cnt = 0, pos = LIST ? rowoff[r] : 0u;
  // plane values at eta index iz: [slot][dx][dy]
  double lo[2][2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int c = 0; c < 2; ++c) lo[a][b][c] = (a ? e1 : e0)[W.cell(ix + b, iy + c, 0)];
  for (int iz = 0; iz < W.dimz; ++iz) {
    int above = 0;
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        for (int c = 0; c < 2; ++c) above += lo[a][b][c] >= W.e_fo;
    if (W.dim == 4) {
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
          for (int c = 0; c < 2; ++c) {
            const double v = (a ? e1 : e0)[W.cell(ix + b, iy + c, iz + 1)];
            above += v >= W.e_fo;
            lo[a][b][c] = v;
          }
    }
    if (above > 0 && above < ncorner) {
      if (LIST) cubes[pos + cnt] = (unsigned)(r * W.dimz + iz);
      ++cnt;
    }
  }
  if (!LIST) rowcnt[r] = cnt;
}

__device__ __forceinline__ void decode_cube(const Window &W, unsigned id, int &itl, int &ix,
                                            int &iy, int &iz) {
  iz = (int)(id % W.dimz);
  unsigned r = id / W.dimz;
  iy = (int)(r % (W.ny - 1));
  r /= (W.ny - 1);
  ix = (int)(r % (W.nx - 1));
  itl = (int)(r / (W.nx - 1));
}

// Corners from the ring (WINDOW) or from an array of cells (test entry).
struct CellSource {
  Window W;
  const unsigned *cubes;
  const double *corners;  // [cell][2^dim] when not a window
  bool window;
};

__global__ __launch_bounds__(64) void fo_cornelius_kernel(CellSource S, long long ncells, int dim,
                                                          double dx0, double dx1, double dx2,
                                                          double dx3, double v0,
                                                          int *__restrict__ ecnt,
                                                          double *__restrict__ elem,
                                                          unsigned *__restrict__ overflow) {
  __shared__ corn::Work W;
  __shared__ double cen[corn::MAXE][corn::D], nrm[corn::MAXE][corn::D];
  const int lane = threadIdx.x;
  const int ncorner = 1 << dim;
  // one block (wave) per cell: cells differ much in cost, and the dispatcher
  // balances single-cell blocks better than a loop over a fixed grid
  {
    const long long cell = blockIdx.x;
    if (cell >= ncells) return;
    if (lane == 0) {
      W.dim = dim;
      W.v0 = v0;
      W.dx[0] = dim == 4 ? dx0 : 1.0;
      W.dx[1] = dim == 4 ? dx1 : dx0;
      W.dx[2] = dim == 4 ? dx2 : dx1;
      W.dx[3] = dim == 4 ? dx3 : dx2;
      W.overflow = 0;
    }
    if (lane < ncorner) {
      double v;
      if (S.window) {
        int itl, ix, iy, iz;
        decode_cube(S.W, S.cubes[cell], itl, ix, iy, iz);
        // corner index over (it, ix, iy[, iz]), last axis fastest
        int a, b, c, d;
        if (dim == 4) {
          a = lane >> 3;
          b = (lane >> 2) & 1;
          c = (lane >> 1) & 1;
          d = lane & 1;
        } else {
          a = lane >> 2;
          b = (lane >> 1) & 1;
          c = lane & 1;
          d = 0;
        }
        v = S.W.e_slot(S.W.start + itl + a)[S.W.cell(ix + b, iy + c, iz + d)];
      } else {
        v = S.corners[cell * ncorner + lane];
      }
      W.v[lane] = v;
    }
    __syncthreads();
    const int n = corn::cell_elements(W, cen, nrm);
    __syncthreads();
    if (lane < n)
      for (int k = 0; k < 4; ++k) {
        elem[(cell * corn::MAXE + lane) * ELEM + k] = cen[lane][k];
        elem[(cell * corn::MAXE + lane) * ELEM + 4 + k] = nrm[lane][k];
      }
    if (lane == 0) {
      ecnt[cell] = n;
      if (W.overflow) atomicAdd(overflow, 1u);
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ emit
struct EmitArgs {
  Window W;
  const unsigned *cubes;
  const unsigned *eoff;
  const double *elem;
  double *rows;
  long long ncubes, nrows;
  double t, t0, dt, dx, dy, dz;
  int n, fofreq, flags, eos;
  EosTable tab;
};

// linearInterp4D / linearInterp3D (FreezeOut.cpp:5-47) with corner values
// a[8 t + 4 x + 2 y + z] (a[4 t + 2 x + y] in 3-D), terms in the reference's
// order, products and sum left to right.
__device__ constexpr unsigned char kTerms4[16] = {0, 8, 4, 2, 1, 12, 10, 9, 6, 5, 3, 14, 13, 11, 7, 15};
__device__ constexpr unsigned char kTerms3[8] = {0, 4, 2, 1, 6, 5, 3, 7};

__device__ double interp4(double x0, double x1, double x2, double x3, const double *a) {
  double r = 0;
  for (int m = 0; m < 16; ++m) {
    const int k = kTerms4[m];
    const double w0 = (k & 8) ? x0 : (1 - x0), w1 = (k & 4) ? x1 : (1 - x1);
    const double w2 = (k & 2) ? x2 : (1 - x2), w3 = (k & 1) ? x3 : (1 - x3);
    const double term = w0 * w1 * w2 * w3 * a[k];
    r = m == 0 ? term : r + term;
  }
  return r;
}

__device__ double interp3(double x0, double x1, double x2, const double *a) {
  double r = 0;
  for (int m = 0; m < 8; ++m) {
    const int k = kTerms3[m];
    const double w0 = (k & 4) ? x0 : (1 - x0), w1 = (k & 2) ? x1 : (1 - x1);
    const double w2 = (k & 1) ? x2 : (1 - x2);
    const double term = w0 * w1 * w2 * a[k];
    r = m == 0 ? term : r + term;
  }
  return r;
}

template <int EOS>
__device__ void eos_T_p(const EosTable &T, double e, double &temp, double &p) {
  temp = eos_T<EOS>(T, e);
  p = eos_p<EOS>(T, e);
}

// One thread per (element, ring variable): the element's cube is the last one
// whose row offset is <= the element's row (eoff is the exclusive scan of the
// element counts), the variable's 2^dim corners are gathered at once and
// interpolated.  Variable 0 also writes the position and normal, variable 3
// (e) also T and P.  Same arithmetic per value as one thread per element.
// Blocks go to the 8 XCDs in turn; the grid is a multiple of 8 and block b
// takes the range (b mod 8) nb/8 + b/8, so each XCD walks a contiguous run of
// the (it, ix, iy, ieta)-ordered elements and neighbouring cube columns,
// which share corner cache lines, meet in the same L2.
template <int DIM>
__global__ __launch_bounds__(256) void fo_emit_kernel(EmitArgs A) {
  const unsigned b = blockIdx.x, nb = gridDim.x;
  const unsigned lb = (b & 7u) * (nb >> 3) + (b >> 3);
  const long long g = (long long)lb * blockDim.x + threadIdx.x;
  if (g >= A.nrows * NVAR) return;
  const long long j = g / NVAR;
  const int v = (int)(g - j * NVAR);
  long long lo = 0, hi = A.ncubes - 1;
  while (lo < hi) {
    const long long mid = (lo + hi + 1) >> 1;
    if ((long long)A.eoff[mid] <= j) lo = mid;
    else hi = mid - 1;
  }
  const long long idx = lo;
  const int i = (int)(j - (long long)A.eoff[idx]);
  const Window &W = A.W;
  int itl, ix, iy, iz;
  decode_cube(W, A.cubes[idx], itl, ix, iy, iz);
  const int it = W.start + itl;
  const double *el = A.elem + (idx * corn::MAXE + i) * ELEM;
  double c[4], m[4], fr[4];
  if constexpr (DIM == 4) {
    for (int k = 0; k < 4; ++k) {
      c[k] = el[k];
      m[k] = el[4 + k];
    }
  } else {
    for (int k = 0; k < 3; ++k) {
      c[k] = el[1 + k];
      m[k] = el[5 + k];
    }
    c[3] = 0;
    m[3] = 0;
  }
  fr[0] = c[0] / A.dt;
  fr[1] = c[1] / A.dx;
  fr[2] = c[2] / A.dy;
  fr[3] = DIM == 4 ? c[3] / A.dz : 0.0;
  double *row = A.rows + (size_t)j * NCOL;
  constexpr int nc = 1 << DIM;
  double a[nc];
#pragma unroll
  for (int k = 0; k < nc; ++k) {
    const int ta = DIM == 4 ? k >> 3 : k >> 2;
    const int xb = DIM == 4 ? (k >> 2) & 1 : (k >> 1) & 1;
    const int yc = DIM == 4 ? (k >> 1) & 1 : k & 1;
    const int zd = DIM == 4 ? k & 1 : 0;
    const double *slot = W.ring + ((long long)W.sl.phys[it + ta] * NVAR + v) * W.nint;
    a[k] = slot[W.cell(ix + xb, iy + yc, iz + zd)];
  }
  double val;
  if constexpr (DIM == 4) val = interp4(fr[0], fr[1], fr[2], fr[3], a);
  else val = interp3(fr[0], fr[1], fr[2], a);
  if (v == 0) {
    // cell tau (HydroPlugin.cpp:304-305)
    double cell_tau;
    if (A.n <= A.fofreq || (A.flags & GVH_FO_EXACT_TAU))
      cell_tau = A.t0 + ((double)(A.n - A.fofreq + (it - 1))) * A.dt;
    else
      cell_tau = A.t0 + ((double)(A.n - A.fofreq + it)) * A.dt;
    const double cell_x = (double)ix * A.dx - (((double)(W.nx - 1)) / 2.0 * A.dx);
    const double cell_y = (double)iy * A.dy - (((double)(W.ny - 1)) / 2.0 * A.dy);
    const double cell_z = (double)iz * A.dz - (((double)(W.nz - 1)) / 2.0 * A.dz);
    row[0] = c[0] + cell_tau;
    row[1] = c[1] + cell_x;
    row[2] = c[2] + cell_y;
    row[3] = DIM == 4 ? c[3] + cell_z : cell_z;
    const double scale = (A.flags & GVH_FO_NORMAL_TAU) ? row[0] : A.t;
    row[4] = scale * m[0];
    row[5] = scale * m[1];
    row[6] = scale * m[2];
    row[7] = DIM == 4 ? scale * m[3] : 0.0;
  }
  row[v < 4 ? 8 + v : 10 + v] = val;
  if (v == 3) {
    double T, P;
    if (A.eos == EOS_CONFORMAL) eos_T_p<EOS_CONFORMAL>(A.tab, val, T, P);
    else if (A.eos == EOS_TABLE) eos_T_p<EOS_TABLE>(A.tab, val, T, P);
    else eos_T_p<EOS_WB>(A.tab, val, T, P);
    row[12] = T;
    row[13] = P;
  }
}

// (Re)allocate a device array for at least `need` items of `per` elements.
template <typename T>
int realloc_dev(T *&p, size_t need, size_t per) {
  if (p) (void)hipFree(p);
  p = nullptr;
  const size_t bytes = need * per * sizeof(T);
  hipError_t e = hipMalloc((void **)&p, bytes ? bytes : sizeof(T));
  if (e != hipSuccess)
    return gvh_fail(GVH_ENOMEM, "freeze-out buffer of %zu bytes: %s", bytes, hipGetErrorString(e));
  return GVH_OK;
}

size_t with_headroom(size_t n) { return n + n / 4 + 1024; }

int ensure_rows_per_window(FoState *f, size_t nrow) {
  if (nrow <= f->row_cap) return GVH_OK;
  const size_t cap = with_headroom(nrow);
  int rc;
  if ((rc = realloc_dev(f->rowcnt, cap, 1)) || (rc = realloc_dev(f->rowoff, cap, 1))) {
    f->row_cap = 0;
    return rc;
  }
  f->row_cap = cap;
  return GVH_OK;
}

int ensure_cubes(FoState *f, size_t ncubes) {
  if (ncubes <= f->cube_cap) return GVH_OK;
  const size_t cap = with_headroom(ncubes);
  int rc;
  if ((rc = realloc_dev(f->cubes, cap, 1)) || (rc = realloc_dev(f->ecnt, cap, 1)) ||
      (rc = realloc_dev(f->eoff, cap, 1)) ||
      (rc = realloc_dev(f->elem, cap, (size_t)corn::MAXE * ELEM))) {
    f->cube_cap = 0;
    return rc;
  }
  f->cube_cap = cap;
  return GVH_OK;
}

int ensure_out_rows(FoState *f, size_t nrows) {
  if (nrows <= f->rows_cap) return GVH_OK;
  const size_t cap = with_headroom(nrows);
  int rc = realloc_dev(f->rows, cap, NCOL);
  f->rows_cap = rc ? 0 : cap;
  return rc;
}

int ensure_scan(FoState *f, size_t n) {
  const size_t need = scan_scratch(n);
  if (need <= f->scan_cap) return GVH_OK;
  const size_t cap = with_headroom(need);
  int rc = realloc_dev(f->scan_tmp, cap, 1);
  f->scan_cap = rc ? 0 : cap;
  return rc;
}

Window make_window(gvh_context *c, FoState *f, int start) {
  Window W{};
  W.ring = f->ring;
  for (int k = 0; k <= f->fofreq; ++k) W.sl.phys[k] = f->slot[k];
  W.nint = f->nint;
  W.nx = c->nx;
  W.ny = c->ny;
  W.nz = c->nz;
  W.dim = f->dim;
  W.dimz = f->dim == 4 ? c->nz - 1 : 1;
  W.start = start;
  W.nit = f->fofreq - start;
  W.e_fo = f->e_fo;
  return W;
}

unsigned grid_for(long long n, int block) { return (unsigned)((n + block - 1) / block); }

}  // namespace

// ================================================================= C ABI

extern "C" int gvh_freezeout_init(gvh_context *c, int fofreq, double e_fo, int flags) {
  return fo_init(c, fofreq, e_fo, flags, 1);
}

// The history of a context with `halves` copies of the fofreq + 1 slots: one
// for a context's own run loop, two for a batched event, whose consecutive
// windows alternate between them so that a window can be searched while the
// next window's steps record into the other half.
namespace gvh {
int fo_init(gvh_context *c, int fofreq, double e_fo, int flags, int halves) {
  if (!c) return gvh_fail(GVH_EINVAL, "context is NULL");
  if (halves != 1 && halves != 2) return gvh_fail(GVH_EINVAL, "halves must be 1 or 2");
  if (fofreq < 3 || fofreq > GVH_FO_MAX_FREQ)
    return gvh_fail(GVH_EINVAL, "fofreq must be in [3, %d], got %d", GVH_FO_MAX_FREQ, fofreq);
  int dim;
  if (c->nx > 1 && c->ny > 1 && c->nz > 1) dim = 4;
  else if (c->nx > 1 && c->ny > 1 && c->nz < 2) dim = 3;
  else
    return gvh_fail(GVH_EINVAL, "freeze-out needs a 2+1D (nx, ny > 1, nz = 1) or 3+1D lattice");
  const long long nint = (long long)c->nx * c->ny * c->nz;
  const long long ncube_max = (long long)(fofreq) * (c->nx - 1) * (c->ny - 1) *
                              (dim == 4 ? c->nz - 1 : 1);
  if (ncube_max * corn::MAXE >= (1ll << 32))
    return gvh_fail(GVH_EINVAL, "lattice too large for 32-bit freeze-out indexing");
  fo_free(c->fo);
  c->fo = nullptr;
  FoState *f = new FoState;
  f->dim = dim;
  f->fofreq = fofreq;
  f->flags = flags;
  f->e_fo = e_fo;
  f->nint = nint;
  f->halves = halves;
  f->slot.resize(fofreq + 1);
  for (int k = 0; k <= fofreq; ++k) f->slot[k] = k;
  const size_t bytes = (size_t)halves * (fofreq + 1) * NVAR * nint * sizeof(double);
  hipError_t e = hipMalloc((void **)&f->ring, bytes);
  if (e == hipSuccess) e = hipMemsetAsync(f->ring, 0, bytes, c->stream);
  if (e == hipSuccess) e = hipMalloc((void **)&f->total, 2 * sizeof(unsigned));
  if (e != hipSuccess) {
    fo_free(f);
    return gvh_fail(GVH_ENOMEM, "freeze-out history of %zu bytes: %s", bytes, hipGetErrorString(e));
  }
  c->fo = f;
  return GVH_OK;
}

// The next window search reads ring half `half` (slots half (fofreq + 1) ..
// half (fofreq + 1) + fofreq), as a batch's record kernel wrote them.
int fo_set_window_half(gvh_context *c, int half) {
  FoState *f = c->fo;
  if (!f || half < 0 || half >= f->halves) return gvh_fail(GVH_EINVAL, "no ring half %d", half);
  for (int k = 0; k <= f->fofreq; ++k) f->slot[k] = half * (f->fofreq + 1) + k;
  return GVH_OK;
}
}  // namespace gvh

// The reference's per-step update of the history (HydroPlugin.cpp:260-269,
// FreezeOut.cpp:49-126) for loop index n >= 1.
extern "C" int gvh_freezeout_record(gvh_context *c, int n) {
  if (!c || !c->fo) return gvh_fail(GVH_EINVAL, "freeze-out not initialised (gvh_freezeout_init)");
  if (!c->initialized) return gvh_fail(GVH_EINVAL, "state not initialised");
  if (n < 1) return gvh_fail(GVH_EINVAL, "loop index n must be >= 1");
  FoState *f = c->fo;
  const int nfo = n % f->fofreq;
  int logical;
  if (nfo == 0) {
    std::swap(f->slot[0], f->slot[f->fofreq]);
    logical = 1;
  } else {
    logical = nfo + 1;
  }
  const bool stale = !(f->flags & GVH_FO_CURRENT_FIELDS) && (c->steps_done & 1);
  const double *u = c->ubuf[stale ? c->iup : c->iu];
  const double *q = c->qbuf[stale ? c->iQ() : c->iq];
  double *dst = f->ring + (size_t)f->slot[logical] * NVAR * f->nint;
  if (c->precision != GVH_FP64)
    fo_record_kernel<float><<<grid_for(f->nint, 256), 256, 0, c->stream>>>(
        c->ebuf[c->ie], u, reinterpret_cast<const float *>(q + 4 * c->G), c->G, c->nx, c->ny, c->nz,
        c->ncx, c->ncy, c->bulk ? 1 : 0, dst, f->nint);
  else
    fo_record_kernel<double><<<grid_for(f->nint, 256), 256, 0, c->stream>>>(
        c->ebuf[c->ie], u, q + 4 * c->G, c->G, c->nx, c->ny, c->nz, c->ncx, c->ncy,
        c->bulk ? 1 : 0, dst, f->nint);
  HIP_TRY(hipGetLastError());
  return GVH_OK;
}

// The window search in three phases, each ending in a small device-to-host
// copy that sizes the next: one context's search runs them back to back
// (gvh_freezeout_find); a batch issues one phase per step for each event
// whose search is in flight, on several streams, once the previous phase
// has finished.
namespace gvh {

// Phase 1: surface cubes per (it, ix, iy) row and their scan; the number of
// cubes is copied to f->htot[0].
int fo_find_count(gvh_context *c, int n, hipStream_t st) {
  FoState *f = c->fo;
  const int start = n <= f->fofreq ? 2 : 0;  // HydroPlugin.cpp:272-275
  Window W = make_window(c, f, start);
  const long long nrow = (long long)(c->nx - 1) * (c->ny - 1) * W.nit;
  int rc;
  if ((rc = ensure_rows_per_window(f, (size_t)nrow)) || (rc = ensure_scan(f, (size_t)nrow)))
    return rc;
  f->n_pending = n;
  f->nrow_pending = nrow;
  f->ncubes = 0;
  f->nrows = 0;
  fo_rows_kernel<false><<<grid_for(nrow, 256), 256, 0, st>>>(W, f->rowcnt, nullptr, nullptr);
  HIP_TRY(hipGetLastError());
  HIP_TRY(scan_with_total(f->rowcnt, f->rowoff, (size_t)nrow, f->scan_tmp, f->total, st));
  HIP_TRY(hipMemcpyAsync(f->htot, f->total, sizeof(unsigned), hipMemcpyDeviceToHost, st));
  return GVH_OK;
}

// Phase 2 (phase 1 complete): the cube list in the reference's order, the
// elements of each cube, and the scan of element counts; the number of rows
// and the overflow count are copied to f->htot.
int fo_find_elements(gvh_context *c, hipStream_t st) {
  FoState *f = c->fo;
  const long long ncubes = f->htot[0];
  f->ncubes = ncubes;
  if (ncubes == 0) {
    f->htot[0] = f->htot[1] = 0;
    return GVH_OK;
  }
  const int start = f->n_pending <= f->fofreq ? 2 : 0;
  Window W = make_window(c, f, start);
  int rc;
  if ((rc = ensure_cubes(f, (size_t)ncubes)) || (rc = ensure_scan(f, (size_t)ncubes))) return rc;
  fo_rows_kernel<true><<<grid_for(f->nrow_pending, 256), 256, 0, st>>>(W, nullptr, f->rowoff,
                                                                        f->cubes);
  HIP_TRY(hipGetLastError());
  CellSource S{W, f->cubes, nullptr, true};
  HIP_TRY(hipMemsetAsync(f->total + 1, 0, sizeof(unsigned), st));
  if (ncubes > 0x7fffffffLL) return gvh_fail(GVH_EINVAL, "%lld surface cubes in one window", ncubes);
  const unsigned grid = (unsigned)ncubes;
  const double dz = c->L.latticeSpacingRapidity;
  const double dt = c->L.latticeSpacingProperTime, dx = c->L.latticeSpacingX,
               dy = c->L.latticeSpacingY;
  if (f->dim == 4)
    fo_cornelius_kernel<<<grid, 64, 0, st>>>(S, ncubes, 4, dt, dx, dy, dz, f->e_fo, f->ecnt,
                                             f->elem, f->total + 1);
  else
    fo_cornelius_kernel<<<grid, 64, 0, st>>>(S, ncubes, 3, dt, dx, dy, 0.0, f->e_fo, f->ecnt,
                                             f->elem, f->total + 1);
  HIP_TRY(hipGetLastError());
  HIP_TRY(scan_with_total((const unsigned *)f->ecnt, f->eoff, (size_t)ncubes, f->scan_tmp,
                          f->total, st));
  HIP_TRY(hipMemcpyAsync(f->htot, f->total, 2 * sizeof(unsigned), hipMemcpyDeviceToHost, st));
  return GVH_OK;
}

// Phase 3 (phase 2 complete): the rows of the elements into f->rows; their
// number in *nelements.  t is the loop's time at this window.
int fo_find_emit(gvh_context *c, double t, hipStream_t st, long long *nelements) {
  FoState *f = c->fo;
  *nelements = 0;
  if (f->ncubes == 0) return GVH_OK;
  if (f->htot[1])
    return gvh_fail(GVH_EINVAL, "%u freeze-out cells exceed %d elements", f->htot[1], corn::MAXE);
  const long long nrows = f->htot[0];
  int rc;
  if ((rc = ensure_out_rows(f, (size_t)nrows))) return rc;
  if (nrows * NVAR > 0xfffff000LL)
    return gvh_fail(GVH_EINVAL, "%lld surface elements in one window", nrows);
  const int start = f->n_pending <= f->fofreq ? 2 : 0;
  EmitArgs A{};
  A.W = make_window(c, f, start);
  A.cubes = f->cubes;
  A.eoff = f->eoff;
  A.elem = f->elem;
  A.rows = f->rows;
  A.ncubes = f->ncubes;
  A.nrows = nrows;
  A.t = t;
  A.t0 = c->Hp.initialProperTimePoint;
  A.dt = c->L.latticeSpacingProperTime;
  A.dx = c->L.latticeSpacingX;
  A.dy = c->L.latticeSpacingY;
  A.dz = c->L.latticeSpacingRapidity;
  A.n = f->n_pending;
  A.fofreq = f->fofreq;
  A.flags = f->flags;
  A.eos = c->eos.kind;
  A.tab = c->tab;
  const unsigned egrid = (grid_for(nrows * NVAR, 256) + 7u) / 8u * 8u;  // multiple of 8 (XCD order)
  if (nrows > 0 && f->dim == 4) fo_emit_kernel<4><<<egrid, 256, 0, st>>>(A);
  else if (nrows > 0) fo_emit_kernel<3><<<egrid, 256, 0, st>>>(A);
  HIP_TRY(hipGetLastError());
  f->nrows = nrows;
  *nelements = nrows;
  return GVH_OK;
}

}  // namespace gvh

extern "C" int gvh_freezeout_find(gvh_context *c, int n, double t, long long *nelements) {
  if (!c || !c->fo) return gvh_fail(GVH_EINVAL, "freeze-out not initialised (gvh_freezeout_init)");
  if (!nelements) return gvh_fail(GVH_EINVAL, "nelements is NULL");
  *nelements = 0;
  hipStream_t st = c->stream;
  int rc;
  if ((rc = fo_find_count(c, n, st))) return rc;
  HIP_TRY(hipStreamSynchronize(st));
  if ((rc = fo_find_elements(c, st))) return rc;
  HIP_TRY(hipStreamSynchronize(st));
  return fo_find_emit(c, t, st, nelements);
}

// Rows of the last search on stream st (asynchronous).
namespace gvh {
int fo_rows_async(gvh_context *c, double *rows, long long nrows, hipStream_t st) {
  if (nrows == 0) return GVH_OK;
  HIP_TRY(hipMemcpyAsync(rows, c->fo->rows, (size_t)nrows * NCOL * sizeof(double),
                         hipMemcpyDeviceToHost, st));
  return GVH_OK;
}
}  // namespace gvh

extern "C" int gvh_freezeout_rows(gvh_context *c, double *rows, long long nrows) {
  if (!c || !c->fo) return gvh_fail(GVH_EINVAL, "freeze-out not initialised (gvh_freezeout_init)");
  if (nrows < 0 || nrows > c->fo->nrows)
    return gvh_fail(GVH_EINVAL, "asked for %lld rows, the last search found %lld", nrows,
                    c->fo->nrows);
  if (nrows == 0) return GVH_OK;
  if (!rows) return gvh_fail(GVH_EINVAL, "rows is NULL");
  HIP_TRY(hipMemcpyAsync(rows, c->fo->rows, (size_t)nrows * NCOL * sizeof(double),
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GVH_OK;
}

extern "C" int gvh_freezeout_slot(gvh_context *c, int logical, int var, double *host_out) {
  if (!c || !c->fo) return gvh_fail(GVH_EINVAL, "freeze-out not initialised (gvh_freezeout_init)");
  FoState *f = c->fo;
  if (logical < 0 || logical > f->fofreq || var < 0 || var >= NVAR || !host_out)
    return gvh_fail(GVH_EINVAL, "gvh_freezeout_slot: bad slot %d or variable %d", logical, var);
  HIP_TRY(hipMemcpyAsync(host_out, f->ring + ((size_t)f->slot[logical] * NVAR + var) * f->nint,
                         (size_t)f->nint * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GVH_OK;
}

extern "C" int gvh_cornelius_cells(int dim, long long ncells, const double *corners,
                                   const double *dx, double v0, int *counts, double *centroids,
                                   double *normals) {
  if (dim != 3 && dim != 4) return gvh_fail(GVH_EINVAL, "dim must be 3 or 4");
  if (ncells < 0 || (ncells > 0 && (!corners || !dx || !counts || !centroids || !normals)))
    return gvh_fail(GVH_EINVAL, "gvh_cornelius_cells: NULL argument");
  if (ncells == 0) return GVH_OK;
  if (ncells > 0x7fffffffLL) return gvh_fail(GVH_EINVAL, "at most 2^31 - 1 cells per call");
  const int ncorner = 1 << dim;
  double *dv = nullptr, *del = nullptr;
  int *dcnt = nullptr;
  unsigned *dov = nullptr;
  hipError_t e = hipMalloc((void **)&dv, (size_t)ncells * ncorner * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void **)&del, (size_t)ncells * corn::MAXE * ELEM * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void **)&dcnt, (size_t)ncells * sizeof(int));
  if (e == hipSuccess) e = hipMalloc((void **)&dov, sizeof(unsigned));
  if (e == hipSuccess) e = hipMemcpy(dv, corners, (size_t)ncells * ncorner * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(dov, 0, sizeof(unsigned));
  if (e == hipSuccess) {
    CellSource S{};
    S.window = false;
    S.corners = dv;
    const unsigned grid = (unsigned)ncells;
    fo_cornelius_kernel<<<grid, 64>>>(S, ncells, dim, dx[0], dx[1], dx[2], dim == 4 ? dx[3] : 0.0,
                                      v0, dcnt, del, dov);
    e = hipGetLastError();
  }
  std::vector<double> el;
  unsigned ov = 0;
  if (e == hipSuccess) {
    el.resize((size_t)ncells * corn::MAXE * ELEM);
    e = hipMemcpy(el.data(), del, el.size() * sizeof(double), hipMemcpyDeviceToHost);
  }
  if (e == hipSuccess) e = hipMemcpy(counts, dcnt, (size_t)ncells * sizeof(int), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(&ov, dov, sizeof ov, hipMemcpyDeviceToHost);
  for (void *p : {(void *)dv, (void *)del, (void *)dcnt, (void *)dov})
    if (p) (void)hipFree(p);
  if (e != hipSuccess) return gvh_fail(GVH_EHIP, "gvh_cornelius_cells: %s", hipGetErrorString(e));
  const int off = dim == 4 ? 0 : 1;
  for (long long i = 0; i < ncells; ++i)
    for (int k = 0; k < counts[i]; ++k)
      for (int j = 0; j < dim; ++j) {
        centroids[(i * corn::MAXE + k) * dim + j] = el[(i * corn::MAXE + k) * ELEM + off + j];
        normals[(i * corn::MAXE + k) * dim + j] = el[(i * corn::MAXE + k) * ELEM + 4 + off + j];
      }
  if (ov) return gvh_fail(GVH_EINVAL, "%u cells exceed %d elements", ov, corn::MAXE);
  return GVH_OK;
}

// ------------------------------------------- cornelius without allocation
// gvh_cornelius_cells_noalloc: the same kernel on device buffers of the
// caller, with its element records and overflow counter in a static device
// array instead of hipMalloc'd memory.  Each GPU holds its own copy of the
// array for the life of the process; host code reaches it only through the
// address hipGetSymbolAddress returns on that device.

namespace {

constexpr long long NA_MAX_CELLS = GVH_CORNELIUS_NOALLOC_MAX_CELLS;
constexpr long long NA_ELEM_DOUBLES = NA_MAX_CELLS * corn::MAXE * ELEM;
// element records, then one 64-bit word holding the overflow counter (the
// tail is padded to 256 bytes)
static __device__ __align__(256) double scratch[NA_ELEM_DOUBLES + 32];

constexpr int NA_MAX_DEVICES = 64;
std::mutex na_mutex;
double *na_scratch[NA_MAX_DEVICES] = {};

// Device address of `scratch` on the current device, looked up once per
// device with that device current.
int noalloc_scratch(double **out) {
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= NA_MAX_DEVICES)
    return gvh_fail(GVH_EINVAL, "gvh_cornelius_cells_noalloc: device %d out of range", dev);
  std::lock_guard<std::mutex> lock(na_mutex);
  if (!na_scratch[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return gvh_fail(GVH_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                      hipGetErrorString(e));
    na_scratch[dev] = static_cast<double *>(p);
  }
  *out = na_scratch[dev];
  return GVH_OK;
}

// Element records -> the caller's centroid and normal arrays ([cell][10][dim]),
// one thread per (cell, element slot); slots past the cell's count are left
// untouched, as gvh_cornelius_cells leaves them.
__global__ __launch_bounds__(256) void cornelius_unpack_kernel(long long ncells, int dim,
                                                               const int *__restrict__ counts,
                                                               const double *__restrict__ elem,
                                                               double *__restrict__ cen,
                                                               double *__restrict__ nrm) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncells * corn::MAXE) return;
  const long long cell = t / corn::MAXE;
  const int k = (int)(t - cell * corn::MAXE);
  if (k >= counts[cell]) return;
  const int off = dim == 4 ? 0 : 1;
  for (int j = 0; j < dim; ++j) {
    cen[t * dim + j] = elem[t * ELEM + off + j];
    nrm[t * dim + j] = elem[t * ELEM + 4 + off + j];
  }
}

}  // namespace

extern "C" int gvh_cornelius_cells_noalloc(int dim, long long ncells, const double *corners,
                                           const double *dx, double v0, int *counts,
                                           double *centroids, double *normals, void *stream) {
  if (dim != 3 && dim != 4) return gvh_fail(GVH_EINVAL, "dim must be 3 or 4");
  if (ncells < 0 || (ncells > 0 && (!corners || !dx || !counts || !centroids || !normals)))
    return gvh_fail(GVH_EINVAL, "gvh_cornelius_cells_noalloc: NULL argument");
  if (ncells > NA_MAX_CELLS)
    return gvh_fail(GVH_EINVAL, "gvh_cornelius_cells_noalloc: %lld cells, at most %lld per call",
                    ncells, NA_MAX_CELLS);
  if (ncells == 0) return GVH_OK;
  double *base = nullptr;
  if (int rc = noalloc_scratch(&base)) return rc;
  double *elem = base;
  unsigned long long *ov_word = reinterpret_cast<unsigned long long *>(base + NA_ELEM_DOUBLES);
  hipStream_t st = static_cast<hipStream_t>(stream);
  // the counter keeps the last call's value: clear it in the stream
  HIP_TRY(launch_fill_u64(ov_word, 1, 0ull, st));
  CellSource S{};
  S.window = false;
  S.corners = corners;
  hipLaunchKernelGGL(fo_cornelius_kernel, dim3((unsigned)ncells), dim3(64), 0, st, S, ncells, dim,
                     dx[0], dx[1], dx[2], dim == 4 ? dx[3] : 0.0, v0, counts, elem,
                     reinterpret_cast<unsigned *>(ov_word));
  HIP_TRY(hipGetLastError());
  const long long nslots = ncells * corn::MAXE;
  hipLaunchKernelGGL(cornelius_unpack_kernel, dim3((unsigned)((nslots + 255) / 256)), dim3(256), 0,
                     st, ncells, dim, counts, elem, centroids, normals);
  HIP_TRY(hipGetLastError());
  unsigned long long ov = 0;
  HIP_TRY(hipMemcpyAsync(&ov, ov_word, sizeof ov, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (ov) return gvh_fail(GVH_EINVAL, "%llu cells exceed %d elements", ov, corn::MAXE);
  return GVH_OK;
}
