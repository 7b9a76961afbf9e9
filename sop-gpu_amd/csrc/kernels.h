// kernels.h -- host-side launchers for the kernels in kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "jobdesc.h"

namespace sop {

// Launch shapes of the LDS step and list kernels that the autotuner picks
// (none of them changes any result): threads of a step unit (its Verlet
// rows), Verlet entries fetched ahead per lane in the step, 64-entry chunks
// in flight per row in the Verlet build, and the row-sort cadence (every
// k-th Verlet rebuild).  split_rows (4, 8 or 16; 0: off) steps the LDS-step
// jobs by k_step_split instead, one workgroup per split_rows rows with each
// row's force terms spread over the workgroup (for a few small jobs), and
// builds their lists one wavefront per row; split_global does the same for
// the jobs above the LDS limit (gathers from global memory).
struct StepShape {
    int unit_rows = kLdsUnit;
    int pair_batch = 6;
    int verlet_chunk = 8;
    int sort_every = kSortEvery;
    int split_rows = 0;
    int split_global = 0;
};
// The autotuner tries split_rows only where the launch has at most this
// many 64-row tiles (the spread pays where one bead per lane leaves CUs idle).
constexpr int kSplitMaxTiles = 128;

// One Langevin step (local step = state.step + kstep) of every job that owns
// a tile.  max_blocks caps the persistent grid (0 = one block per 4 tiles).
// Jobs flagged lds_step are stepped by the LDS kernel over `units` (first
// segment, segment count) of `segs` (job slot, first row, rows, LDS offset
// in float4; lds_beads = the largest LDS use of a unit), the others by the
// tile kernel (launched when any_global is set).  With shape.split_rows set
// the LDS-step jobs are stepped by k_step_split over the tiles instead.
hipError_t launch_step(const DevJob* jobs, const DevJobState* st, const int2* tiles, int ntiles, int any_global,
                       const int2* units, const int4* segs, int nunits, int lds_beads, int kstep, int max_blocks,
                       hipStream_t stream, const StepShape& shape);
// List rebuilds for jobs whose cadence is due (force != 0: rebuild all).
// cell_words: bitmap words (ceil(n/64)) of the largest job with a cell grid,
// 0 to build every job's possible list by the all-pairs scan.
// sunits/segs/nsunits/lds_beads: the LDS step plan (see launch_step), whose
// jobs get their possible lists scanned from LDS-staged coordinates
// (nsunits = 0: all jobs by the tile kernels); any_global: some job is
// outside that plan (the tile kernel is launched only then).
hipError_t launch_possible(const DevJob* jobs, DevJobState* st, const int2* tiles, int ntiles, int njobs,
                           int cell_words, int any_global, const int2* sunits, const int4* segs, int nsunits,
                           int lds_beads, int force, hipStream_t stream, const StepShape& shape);
// Largest cell_words launch_possible accepts (the LDS bitmap of one wavefront).
constexpr int kMaxCellWords = 8192;
// any_idx32: some job of the launch stores its lists as int32 (n >= 65536).
// units (job, first row) of the jobs whose rows are re-sorted first;
// sunits/segs/nsunits/lds_beads: the LDS step plan (see launch_step), whose
// jobs get their lists built from LDS-staged coordinates (nsunits = 0: all
// jobs by the tile kernel); any_global: some job is outside that plan.
hipError_t launch_verlet(const DevJob* jobs, DevJobState* st, const int2* tiles, int ntiles, int any_idx32,
                         int any_global, const int2* units, int nunits, const int2* sunits, const int4* segs,
                         int nsunits, int lds_beads, int force, hipStream_t stream, const StepShape& shape);
// Per-job updaters at a block start, and the step counter advance.
hipError_t launch_job_update(const DevJob* jobs, DevJobState* st, int njobs, hipStream_t stream, int mask = 7);
hipError_t launch_job_advance(const DevJob* jobs, DevJobState* st, int njobs, int nsteps, hipStream_t stream);
// Output / parity hooks.
hipError_t launch_energies(const DevJob* jobs, const DevJobState* st, int job, int n, hipStream_t stream);
// Forces only (the step kernels' force terms, no state change) of the jobs
// of the given tiles / LDS plan, into out[bead] (one job per launch).
hipError_t launch_forces(const DevJob* jobs, const DevJobState* st, const int2* tiles, int ntiles, int any_global,
                         const int2* units, const int4* segs, int nunits, int lds_beads, float4* out,
                         hipStream_t stream, const StepShape& shape);
hipError_t launch_rng_probe(uint4* state, int n, int count, int gaussian, float* out, hipStream_t stream);
// Counts (into *bad, device) the b bit patterns in [lo, hi) times numerators
// a[0..na) (device) whose short division differs from the IEEE quotient.
hipError_t launch_division_check(unsigned lo, unsigned hi, const float* a, int na, unsigned long long* bad,
                                 hipStream_t stream);
const void* step_kernel_symbol();
hipError_t launch_copy(const float4* src, float4* dst, size_t n, hipStream_t stream);

}  // namespace sop
