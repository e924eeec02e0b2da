// gvh_batch_kernels.hip -- kernels of a batch of events (gvh_batch.hip):
// the stage kernel over a tile map, the history record with the count of
// cells above the freeze-out energy density, the termination check and the
// per-step advance.
#include "gvh_stage_body.h"

namespace gvh {

// ------------------------------------------------------------ batches
// Many events of one equation of state, bulk setting and precision in one
// launch: blocks loop over the tiles of a tile map (tile -> slot), and each
// tile is one block's worth of cells of one event, cut exactly as that event
// alone would be cut, so every event gets the arithmetic of a launch of its
// own.  The stage pointers and times come from the slot's roles and step
// count.  One block per tile (no tile loop): the slot is read before any
// store of the block, so its fields stay in scalar registers; a loop over
// tiles made the compiler reload them into vector registers, which cost a
// wave per SIMD.
template <int EOS, bool BULK, bool STAGE2, bool HAS_Z, bool KEEP, int LOCK, typename R,
          typename PS>
__global__ __launch_bounds__(BLOCK, stage_waves_per_simd<R>()) void kt_batch_kernel(
    const BatchSlot *__restrict__ slots, const int *__restrict__ tile_slot, int begin) {
  __shared__ double acc_lds[15 * BLOCK];
  const int t = begin + (int)blockIdx.x;
  const BatchSlot &S = slots[tile_slot[t]];
  if (S.status != SLOT_RUN) return;
  StageArgs A = S.A;
  stage_pointers(A, S.B, S.R, STAGE2 ? 2 : 1);
  stage_times(A, S.t0 + S.steps * S.A.dt, STAGE2 ? 2 : 1);
  const CellPos c = cell_of_tile(A, (unsigned)(t - S.tile_first), (unsigned)S.ntiles);
  if (c.valid)
    stage_cell<EOS, BULK, STAGE2, HAS_Z, true, KEEP, LOCK, R, PS>(A, c, acc_lds + threadIdx.x);
}

// The run loop's per-step history update (FreezeOut.cpp:49-126) for every
// running event at loop index n = steps + 1, with the ring's slot 0 <- slot
// FOFREQ done as a copy, and the count of interior cells with e > e_fo for
// the termination rule (HydroPlugin.cpp:396-407).  u and pi are the previous
// step's after an odd number of steps unless GVH_FO_CURRENT_FIELDS.
template <typename PS>
__global__ __launch_bounds__(BLOCK) void batch_record_kernel(BatchSlot *__restrict__ slots,
                                                             const int *__restrict__ tile_slot) {
  const int t = blockIdx.x;
  const int sl = tile_slot[t];
  BatchSlot &S = slots[sl];
  if (S.status != SLOT_RUN || S.steps >= S.max_steps) return;
  const CellPos c = cell_of_tile(S.A, (unsigned)(t - S.tile_first), (unsigned)S.ntiles);
  const long long G = S.A.G;
  const double *e = S.B.e[S.R.ie];
  bool above = false;
  if (c.valid) {
    const unsigned s = c.s;
    above = e[s] > S.e_fo;
    if (S.ring) {
      const bool stale = !(S.fo_flags & GVH_FO_CURRENT_FIELDS) && (S.steps & 1);
      const double *u = S.B.u[stale ? S.R.iup : S.R.iu];
      const PS *qP = reinterpret_cast<const PS *>(S.B.q[stale ? 1 - S.R.iq : S.R.iq] + 4 * G);
      const long long nint = S.nint;
      const long long cell =
          (c.i - NG) + (long long)S.A.nx * ((c.j - NG) + (long long)S.A.ny * (c.k - NG));
      const int n = S.steps + 1, nfo = n % S.fofreq;
      // window m = n / fofreq records into half m mod 2 when the ring has two
      // halves; its slot 0 is the previous window's last slot
      const int m = n / S.fofreq;
      const long long span = (long long)(S.fofreq + 1) * 10 * nint;
      double *ring = S.ring + (S.halves == 2 && (m & 1) ? span : 0);
      const double *prev = S.ring + (S.halves == 2 && m > 0 && ((m - 1) & 1) ? span : 0);
      if (nfo == 0)
        for (int v = 0; v < 10; ++v)
          ring[(long long)v * nint + cell] = prev[((long long)S.fofreq * 10 + v) * nint + cell];
      double *dst = ring + (long long)(nfo == 0 ? 1 : nfo + 1) * 10 * nint + cell;
      dst[0 * nint] = u[1 * G + s];
      dst[1 * nint] = u[2 * G + s];
      dst[2 * nint] = u[3 * G + s];
      dst[3 * nint] = e[s];
      dst[4 * nint] = (double)qP[(GVH_PIXX - 4) * G + s];
      dst[5 * nint] = (double)qP[(GVH_PIXY - 4) * G + s];
      dst[6 * nint] = (double)qP[(GVH_PIXN - 4) * G + s];
      dst[7 * nint] = (double)qP[(GVH_PIYY - 4) * G + s];
      dst[8 * nint] = (double)qP[(GVH_PIYN - 4) * G + s];
      dst[9 * nint] = S.bulk ? (double)qP[(GVH_PI - 4) * G + s] : 0.0;
    }
  }
  // one atomic per wave
  const unsigned long long m = __ballot(above);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(&S.above, (unsigned long long)__popcll(m));
}

// Termination before the step (HydroPlugin.cpp:239, 396-413): an event ends
// after nt + 1 steps, or once FOFREQ + 1 recorded states had no cell above
// e_fo.  An event with a Newton iteration budget ends, with its error flag
// set, after a step in which a root-find reached the budget (the reference
// has no budget: it iterates up to MAX_ITERS, EnergyMomentumTensor.cu:22-44).
__global__ void batch_check_kernel(BatchSlot *__restrict__ slots, int nslots) {
  const int sl = blockIdx.x * blockDim.x + threadIdx.x;
  if (sl >= nslots) return;
  BatchSlot &S = slots[sl];
  if (S.status != SLOT_RUN) return;
  if (S.A.max_iters < MAX_ITERS && S.A.diag[DIAG_NOT_CONVERGED] != 0) {
    S.status = SLOT_DONE;
    S.error = 1;
    return;
  }
  if (S.steps >= S.max_steps) {
    S.status = SLOT_DONE;
    return;
  }
  if (S.above == 0) S.frozen += 1;
  if (S.frozen >= S.fofreq + 1) {
    S.status = SLOT_DONE;
    S.frozen_stop = 1;
  }
}

// After the two stages: new buffer roles, one more step, counter cleared.
__global__ void batch_advance_kernel(BatchSlot *__restrict__ slots, int nslots) {
  const int sl = blockIdx.x * blockDim.x + threadIdx.x;
  if (sl >= nslots) return;
  BatchSlot &S = slots[sl];
  if (S.status == SLOT_RUN) {
    Roles R = S.R;
    rotate_roles(R);
    S.R = R;
    S.steps += 1;
  }
  S.above = 0;
}

template <int EOS, bool BULK, bool STAGE2, bool HAS_Z, bool LOCK, bool MIXED>
static void launch_b(const BatchSlot *slots, const int *tile_slot, int begin, int end,
                     hipStream_t st) {
  if constexpr (MIXED)
    hipLaunchKernelGGL((kt_batch_kernel<EOS, BULK, STAGE2, HAS_Z, true, LOCK ? 1 : 0, float, float>),
                       dim3(end - begin), dim3(BLOCK), 0, st, slots, tile_slot, begin);
  else
    hipLaunchKernelGGL((kt_batch_kernel<EOS, BULK, STAGE2, HAS_Z, true, LOCK ? 1 : 0, double, double>),
                       dim3(end - begin), dim3(BLOCK), 0, st, slots, tile_slot, begin);
}

template <int EOS, bool BULK, bool STAGE2, bool HAS_Z, bool LOCK>
static void launch_bm(const BatchSlot *slots, const int *tile_slot, int begin, int end,
                      bool mixed, hipStream_t st) {
  if (mixed) launch_b<EOS, BULK, STAGE2, HAS_Z, LOCK, true>(slots, tile_slot, begin, end, st);
  else launch_b<EOS, BULK, STAGE2, HAS_Z, LOCK, false>(slots, tile_slot, begin, end, st);
}

template <int EOS, bool BULK, bool STAGE2>
static void launch_bz(const BatchSlot *slots, const int *tile_slot, int begin, int end,
                      bool has_z, bool lock, bool mixed, hipStream_t st) {
  if (has_z) {
    if (lock) launch_bm<EOS, BULK, STAGE2, true, true>(slots, tile_slot, begin, end, mixed, st);
    else launch_bm<EOS, BULK, STAGE2, true, false>(slots, tile_slot, begin, end, mixed, st);
  } else {
    if (lock) launch_bm<EOS, BULK, STAGE2, false, true>(slots, tile_slot, begin, end, mixed, st);
    else launch_bm<EOS, BULK, STAGE2, false, false>(slots, tile_slot, begin, end, mixed, st);
  }
}

template <int EOS>
static void launch_be(const BatchSlot *slots, const int *tile_slot, int begin, int end,
                      bool bulk, bool stage2, bool has_z, bool lock, bool mixed, hipStream_t st) {
  if (bulk) {
    if (stage2) launch_bz<EOS, true, true>(slots, tile_slot, begin, end, has_z, lock, mixed, st);
    else launch_bz<EOS, true, false>(slots, tile_slot, begin, end, has_z, lock, mixed, st);
  } else {
    if (stage2) launch_bz<EOS, false, true>(slots, tile_slot, begin, end, has_z, lock, mixed, st);
    else launch_bz<EOS, false, false>(slots, tile_slot, begin, end, has_z, lock, mixed, st);
  }
}

// One RK stage of every running event with tiles [begin, end) of the tile
// map; those events share the equation of state, bulk setting,
// dimensionality, arrangement and precision.
hipError_t launch_batch_stage(const BatchSlot *slots, const int *tile_slot, int begin, int end,
                              int eos, bool bulk, bool stage2, bool has_z, bool lock,
                              bool mixed, hipStream_t st) {
  switch (eos) {
    case EOS_CONFORMAL:
      launch_be<EOS_CONFORMAL>(slots, tile_slot, begin, end, bulk, stage2, has_z, lock, mixed,
                               st);
      break;
    case EOS_WB:
      launch_be<EOS_WB>(slots, tile_slot, begin, end, bulk, stage2, has_z, lock, mixed, st);
      break;
    default:
      launch_be<EOS_TABLE>(slots, tile_slot, begin, end, bulk, stage2, has_z, lock, mixed, st);
  }
  return hipGetLastError();
}

// One block per tile of the tile map.
hipError_t launch_batch_record(BatchSlot *slots, const int *tile_slot, int ntiles, bool mixed,
                               hipStream_t st) {
  if (mixed)
    hipLaunchKernelGGL(batch_record_kernel<float>, dim3(ntiles), dim3(BLOCK), 0, st, slots,
                       tile_slot);
  else
    hipLaunchKernelGGL(batch_record_kernel<double>, dim3(ntiles), dim3(BLOCK), 0, st, slots,
                       tile_slot);
  return hipGetLastError();
}

hipError_t launch_batch_check(BatchSlot *slots, int nslots, hipStream_t st) {
  hipLaunchKernelGGL(batch_check_kernel, dim3((nslots + 63) / 64), dim3(64), 0, st, slots, nslots);
  return hipGetLastError();
}

hipError_t launch_batch_advance(BatchSlot *slots, int nslots, hipStream_t st) {
  hipLaunchKernelGGL(batch_advance_kernel, dim3((nslots + 63) / 64), dim3(64), 0, st, slots,
                     nslots);
  return hipGetLastError();
}

}  // namespace gvh
