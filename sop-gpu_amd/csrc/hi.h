// hi.h -- launchers of the hydrodynamics kernels (hi.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "jobdesc.h"

namespace sop {

enum { SOP_CHOL_BF16X3 = 0, SOP_CHOL_F32 = 1 };

// Width of the factorisation's outer panel: the split-BF16 panel buffers
// hold two panels of mp x kCholPanel elements each (consecutive panels
// alternate: a panel's trailing update, and the next panel's steps, still
// read its copy while the next panel is factored).
constexpr int kCholPanel = 512;
// Step buffers of the 64-column steps: the inverses of two consecutive 64x64
// diagonal blocks (FP64, or FP32 in the first half of the slot in the
// split-BF16 mode), then two FP32 64x64 blocks that carry a diagonal block,
// updated by all but its last 64 columns, from one step to the next.
constexpr size_t kCholLinvBytes = (size_t)2 * 64 * 64 * sizeof(double) + (size_t)2 * 64 * 64 * sizeof(float);
// Column blocks left for the next step (two consecutive steps): mp x 64 FP32 each.
inline size_t chol_bbuf_bytes(size_t mp) { return (size_t)2 * mp * 64 * sizeof(float); }
// One split-BF16 half (hi or lo) of the two panel buffers.
inline size_t chol_panel_bytes(size_t mp) { return (size_t)2 * mp * kCholPanel * sizeof(unsigned short); }

// Device workspace of one blocked factorisation.
struct CholWork {
    double* linv;          // kCholLinvBytes: two inverse slots (FP64-sized), then two FP32 diagonal blocks
    float* bbuf;           // chol_bbuf_bytes(mp)
    unsigned short* phi;   // chol_panel_bytes(mp): split-BF16 hi of the outer panels
    unsigned short* plo;   // chol_panel_bytes(mp): lo
};

// Device workspace of one exact-HI job's blocked factorisation.
struct HiExactWork {
    float* M;
    size_t ld;
    int mp;                // padded order (multiple of 64)
    int mode;              // SOP_CHOL_*
    CholWork work;
    int chol_freq;         // refactorisation interval in steps
    hipStream_t aux;       // stream of the wide trailing updates (or null)
    hipEvent_t ev_f, ev_t2[2];
};

hipError_t launch_hi_tea_update(const DevJob* jobs, DevJobState* st, const int2* tiles, int ntiles, int njobs,
                                hipStream_t stream);
hipError_t launch_hi_tea_step(const DevJob* jobs, const DevJobState* st, const int2* tiles, int ntiles, int kstep,
                              hipStream_t stream);
// D and D.F for the step, the factorisation (when w is given: the caller
// knows the step is a refactorisation step), then the noise and the Euler
// step, for the exact-HI job the tiles belong to.
hipError_t launch_chol_step(const DevJob* jobs, const DevJobState* st, const int2* tiles, int ntiles, int kstep,
                            const HiExactWork* w, hipStream_t stream);
// Standalone factorisation of M (mp x mp, row-major, lower triangle used).
// aux, ev_f, ev_t2a, ev_t2b (optional): a second stream and three events
// for the overlapped trailing updates; the factor is complete in `stream`
// order on return.
hipError_t launch_chol_factor(float* M, size_t ld, int mp, int mode, const CholWork& w, hipStream_t stream,
                              hipStream_t aux = nullptr, hipEvent_t ev_f = nullptr, hipEvent_t ev_t2a = nullptr,
                              hipEvent_t ev_t2b = nullptr);

}  // namespace sop
