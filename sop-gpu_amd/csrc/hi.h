// hi.h -- launchers of the hydrodynamics kernels (hi.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "jobdesc.h"

namespace sop {

enum { SOP_CHOL_BF16X3 = 0, SOP_CHOL_F32 = 1 };

// Width of the factorisation's outer panel: the split-BF16 panel buffers
// hold two panels of mp x kCholPanel elements each (consecutive panels
// alternate, so a panel's trailing update may still read its copy while
// the next panel is factored).
constexpr int kCholPanel = 256;
// FP64 inverses of an outer panel's 64x64 diagonal blocks.
constexpr size_t kCholLinvBytes = (size_t)(kCholPanel / 64) * 64 * 64 * sizeof(double);

// Device workspace of one exact-HI job's blocked factorisation.
struct HiExactWork {
    float* M;
    size_t ld;
    int mp;                // padded order (multiple of 64)
    int mode;              // SOP_CHOL_*
    double* linv;          // kCholLinvBytes: inverses of the current panel's diagonal blocks
    unsigned short* phi;   // split-BF16 outer panels, two buffers of mp x kCholPanel
    unsigned short* plo;
    int chol_freq;         // refactorisation interval in steps
    hipStream_t aux;       // lookahead stream of the trailing updates (or null)
    hipEvent_t ev_f, ev_t2;
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
// aux, ev_f, ev_t2 (optional): a second stream and two events for the
// lookahead schedule; the factor is complete in `stream` order on return.
hipError_t launch_chol_factor(float* M, size_t ld, int mp, int mode, double* linv, unsigned short* phi,
                              unsigned short* plo, const int* flag, hipStream_t stream, hipStream_t aux = nullptr,
                              hipEvent_t ev_f = nullptr, hipEvent_t ev_t2 = nullptr);

}  // namespace sop
