// gvh_stage.hip -- launchers of the one-event stage kernels (kt_flux_kernel,
// kt_source_kernel in gvh_stage_body.h), the termination count and a fill
// kernel.
#include "gvh_stage_body.h"

namespace gvh {

// ------------------------------------------------------------ launch
// Kernel arrangement of a stage (StageArgs::variant): 0 fused kernel with
// the inversion states kept in registers, 1 fused with states rebuilt one at
// a time, 2 separate source kernel + flux kernel (states rebuilt), 3 as 0
// with the four root-finds of a direction in lockstep, 4 as 3 with the
// states rebuilt for the flux after the root-finds.  The mixed form
// (FP32 stencil, limiter, flux and sources, FP32 pi/Pi storage, FP32
// half-site root-finds for the tabulated and conformal EoS; FP64
// T^{tau mu}, accumulator, the cell's own root-find and its EoS) has
// arrangements 0, 3 and 4; others run as 0.  With root64 (precision
// GVH_MIXED_F64_ROOTS) the mixed form keeps the half-site root-finds in
// FP64: arrangement 3 runs its lockstep in FP64, and arrangement 0's
// one-at-a-time root-finds are FP64 in every form.  Arrangement 5 is 3 in
// blocks of four rows that stage the y stencil in LDS
// (kt_flux_kernel_ystaged); it is built for the tabulated EoS with bulk in
// FP64 and GVH_MIXED, needs a strip height that is a multiple of 4, and runs
// as 3 otherwise.
template <int EOS, bool BULK, bool STAGE2, bool HAS_Z, bool MIXED>
static hipError_t launch_z(const StageArgs &A, bool root64, hipStream_t st) {
  const long long ncell = (long long)A.nx * A.strip * A.nstrips * A.nz;
  const int block = BLOCK;
  const int grid = (int)(((ncell + block - 1) / block + 7) / 8 * 8);  // multiple of 8 (XCD order)
  if constexpr (EOS == EOS_TABLE && BULK) {
    if (A.variant == 5 && !root64 && A.strip % YROWS == 0 && A.ny >= YROWS) {
      const long long nb = (ncell + YROWS * block - 1) / (YROWS * block);
      const int grid4 = (int)((nb + 7) / 8 * 8);
      using R = typename std::conditional<MIXED, float, double>::type;
      hipLaunchKernelGGL((kt_flux_kernel_ystaged<EOS, BULK, STAGE2, HAS_Z, 1, R, R>),
                         dim3(grid4), dim3(YROWS * block), 0, st, A);
      return hipGetLastError();
    }
  }
  if constexpr (MIXED) {
    // FP32 stencil, limiter, flux and sources; FP32 pi/Pi storage
    if (root64 && A.variant != 0)
      hipLaunchKernelGGL((kt_flux_kernel<EOS, BULK, STAGE2, HAS_Z, true, true, 3, float, float>),
                         dim3(grid), dim3(block), 0, st, A);
    else if (A.variant == 3 || A.variant == 5)
      hipLaunchKernelGGL((kt_flux_kernel<EOS, BULK, STAGE2, HAS_Z, true, true, 1, float, float>),
                         dim3(grid), dim3(block), 0, st, A);
    else if (A.variant == 4)
      hipLaunchKernelGGL((kt_flux_kernel<EOS, BULK, STAGE2, HAS_Z, true, true, 2, float, float>),
                         dim3(grid), dim3(block), 0, st, A);
    else
      hipLaunchKernelGGL((kt_flux_kernel<EOS, BULK, STAGE2, HAS_Z, true, true, 0, float, float>),
                         dim3(grid), dim3(block), 0, st, A);
    return hipGetLastError();
  } else {
    switch (A.variant) {
      case 0:
        hipLaunchKernelGGL(
            (kt_flux_kernel<EOS, BULK, STAGE2, HAS_Z, true, true, 0, double, double>),
            dim3(grid), dim3(block), 0, st, A);
        break;
      case 1:
        hipLaunchKernelGGL(
            (kt_flux_kernel<EOS, BULK, STAGE2, HAS_Z, true, false, 0, double, double>),
            dim3(grid), dim3(block), 0, st, A);
        break;
      case 3:
      case 5:
        hipLaunchKernelGGL(
            (kt_flux_kernel<EOS, BULK, STAGE2, HAS_Z, true, true, 1, double, double>),
            dim3(grid), dim3(block), 0, st, A);
        break;
      case 4:
        hipLaunchKernelGGL(
            (kt_flux_kernel<EOS, BULK, STAGE2, HAS_Z, true, true, 2, double, double>),
            dim3(grid), dim3(block), 0, st, A);
        break;
      default:
        hipLaunchKernelGGL((kt_source_kernel<EOS, BULK, HAS_Z>), dim3(grid), dim3(block), 0, st,
                           A);
        hipLaunchKernelGGL(
            (kt_flux_kernel<EOS, BULK, STAGE2, HAS_Z, false, false, 0, double, double>),
            dim3(grid), dim3(block), 0, st, A);
    }
    return hipGetLastError();
  }
}

template <int EOS, bool BULK, bool STAGE2, bool MIXED>
static hipError_t launch_t(const StageArgs &A, bool has_z, bool root64, hipStream_t st) {
  return has_z ? launch_z<EOS, BULK, STAGE2, true, MIXED>(A, root64, st)
               : launch_z<EOS, BULK, STAGE2, false, MIXED>(A, root64, st);
}

template <int EOS, bool MIXED>
static hipError_t launch_e(const StageArgs &A, bool bulk, bool stage2, bool has_z, bool root64,
                           hipStream_t st) {
  if (bulk)
    return stage2 ? launch_t<EOS, true, true, MIXED>(A, has_z, root64, st)
                  : launch_t<EOS, true, false, MIXED>(A, has_z, root64, st);
  return stage2 ? launch_t<EOS, false, true, MIXED>(A, has_z, root64, st)
                : launch_t<EOS, false, false, MIXED>(A, has_z, root64, st);
}

template <bool MIXED>
static hipError_t launch_p(const StageArgs &A, int eos, bool bulk, bool stage2, bool root64,
                           hipStream_t st) {
  const bool has_z = A.nz > 1;
  switch (eos) {
    case EOS_CONFORMAL: return launch_e<EOS_CONFORMAL, MIXED>(A, bulk, stage2, has_z, root64, st);
    case EOS_WB: return launch_e<EOS_WB, MIXED>(A, bulk, stage2, has_z, root64, st);
    default: return launch_e<EOS_TABLE, MIXED>(A, bulk, stage2, has_z, root64, st);
  }
}

// precision: GVH_FP64, GVH_MIXED or GVH_MIXED_F64_ROOTS
hipError_t launch_stage(const StageArgs &A, int eos, bool bulk, bool stage2, int precision,
                        hipStream_t st) {
  if (precision == GVH_FP64) return launch_p<false>(A, eos, bulk, stage2, false, st);
  return launch_p<true>(A, eos, bulk, stage2, precision == GVH_MIXED_F64_ROOTS, st);
}

// Termination scan of the run loop (HydroPlugin.cpp:396-407): count interior
// cells with e > threshold; one atomic per block.
__global__ __launch_bounds__(256) void count_above_kernel(const double *__restrict__ e, int nx,
                                                          int ny, int nz, int ncx, int ncy,
                                                          double thr,
                                                          unsigned long long *__restrict__ out) {
  __shared__ unsigned int wsum[4];
  const int nint = nx * ny * nz;
  unsigned int c = 0;
  for (int tid = blockIdx.x * blockDim.x + threadIdx.x; tid < nint; tid += gridDim.x * blockDim.x) {
    const int i = tid % nx + NG, r = tid / nx, j = r % ny + NG, k = r / ny + NG;
    c += e[i + (long long)ncx * (j + (long long)ncy * k)] > thr ? 1u : 0u;
  }
  // wavefront reduction (64 lanes) then across the 4 waves of the block
  for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) wsum[w] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long tot = 0;
    for (int m = 0; m < (int)(blockDim.x >> 6); ++m) tot += wsum[m];
    if (tot) atomicAdd(out, tot);
  }
}

hipError_t launch_count_above(const double *e, int nx, int ny, int nz, int ncx, int ncy, double thr,
                              unsigned long long *out, hipStream_t st) {
  const int nint = nx * ny * nz;
  int grid = (nint + 255) / 256;
  if (grid > 2048) grid = 2048;
  hipLaunchKernelGGL(count_above_kernel, dim3(grid), dim3(256), 0, st, e, nx, ny, nz, ncx, ncy, thr,
                     out);
  return hipGetLastError();
}

// Fill kernel for clearing small device buffers inside a stream without
// hipMemsetAsync.
__global__ void fill_u64_kernel(unsigned long long *p, int n, unsigned long long v) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

hipError_t launch_fill_u64(unsigned long long *p, int n, unsigned long long v, hipStream_t st) {
  hipLaunchKernelGGL(fill_u64_kernel, dim3((n + 63) / 64), dim3(64), 0, st, p, n, v);
  return hipGetLastError();
}

}  // namespace gvh
