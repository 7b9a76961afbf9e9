// jobdesc.h -- device-side description of one simulated system ("job").
//
// Every job owns its arrays inside the engine's HBM arena.  Neighbour
// lists are slot-major as in the reference (entry k of bead i at
// [k*n + i], src/Potentials/covalent.cu:85-96) so that one lane per bead
// reads consecutive words, and they hold job-local bead indices.
// Coordinates are double-buffered: the coordinates at local step s live in
// coord[s & 1]; step s reads that buffer and writes the other, so one
// kernel per step needs no grid-wide barrier.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sop {

enum JobFlags : int {
    JF_MINIMIZATION = 1,
    JF_PULLING = 2,      // extForces in use (fixed / pulled beads)
    JF_PULL_UPDATER = 4, // force-ramp cantilever updated every stride
    JF_INDENTATION = 8,
    JF_HEATING = 16,
    JF_EXTRA_FORCE = 32, // forces accumulated by non-built-in potentials
    JF_HI = 64,          // hydrodynamic interactions (bdhitea)
    JF_HI_EXACT = 128,   // Cholesky path instead of TEA
    JF_HI_LISTED = 256,  // TEA over the covalent/native/Verlet lists only
    JF_HI_FAITHFUL = 512 // Cholesky noise applied block-symmetrised, as the reference reads L
};

struct DevJob {
    int n;
    int flags;
    // arrays (device pointers into the arena)
    float4* coord[2];
    uint4* rng;
    float* T;
    float4* ext;    // x,y,z force, w: 0 free, 1 fixed, 2 pulled
    unsigned char* bflag;  // ext[i].w as one byte, read by every bead's step (ext only for pulled beads)
    float4* tipF;   // per-bead reaction on the indentation tip
    float4* fextra; // optional extra forces (custom potentials)
    int* cov_cnt;
    int* cov_j;
    unsigned short* cov_j16;  // the same entries as 16-bit indices (n < 65,536; else null), read by the step
    float* cov_r0;
    int* nat_cnt;
    int* nat_j;
    unsigned short* nat_j16;  // the same entries as 16-bit indices (n < 65,536; else null), read by the step
    float2* nat_par; // (r0^2, -12 eh / r0^2)
    // Verlet and possible lists hold job-local bead indices as uint16 when the
    // job has fewer than 65536 beads (idx16), as int32 otherwise
    // Verlet list rows: row p holds the list of bead perm[p] (slot-major:
    // entry k of row p at [k*pair_ld + p]) and the wavefront that steps rows
    // 64g..64g+63 steps those beads.  For jobs with sort_rows set (stepped
    // by k_step_lds, not hydrodynamic) the beads of each kLdsUnit-row unit
    // are re-sorted by list length at every 20th Verlet rebuild, so a
    // wavefront's lists are of about equal length; elsewhere perm is the
    // identity.
    int* perm;
    int sort_rows;
    // The bonded lists in Verlet-row order, for jobs whose rows are sorted
    // (refreshed with every sort): counts per row (covalent | native << 16)
    // and slot-major entries [k*pair_ld + r] of bead perm[r] (else null).
    int* rb_cnt;
    unsigned short* rb_cov_j;
    float* rb_cov_r0;
    unsigned short* rb_nat_j;
    float2* rb_nat_par;
    int* pair_cnt;  // per row
    void* pair_j;
    int* pp_cnt;
    void* pp_j;     // bead-major: entry k of bead i at [i*pp_ld + k]
    int* excl_cnt;  // sorted covalent+native partners (possible-pair exclusions)
    int* excl_j;
    // cell grid of the possible-pairs build, for jobs of kCellMinBeads
    // beads and more (null otherwise): per-cell counts and exclusive starts,
    // the beads grouped by cell, each bead's cell and slot within it
    int* cell_cnt;
    int* cell_start;
    int* cell_beads;
    int* bead_cell;
    int* bead_slot;
    int cell_cap;
    int lds_step;   // stepped by k_step_lds (coordinates staged in LDS)
    int* pulled;    // bead indices of pulled beads
    float* energies; // 3n per-bead energies (output cadence)
    int cov_cap, nat_cap, pair_cap, pp_cap, excl_cap;
    int pair_ld;    // Verlet slot stride: n rounded up to 64 rows (entry k of row r at [k*pair_ld + r])
    int pp_ld;      // possible-list row stride: pp_cap rounded up to 64 entries (entry k of bead i at [i*pp_ld + k])
    int idx16;
    int npulled;
    // force-field constants, formed as the reference's constructors do
    float ks, R2;                     // FENE: kspring_cov, R_limit^2
    float a2, m6, rc2, el;            // repulsive LJ: a^2, -6 el/a^2, cutoff^2
    float pl_cut, pp_thr;             // Verlet cutoff, possible-pairs threshold
    // largest float d2 with sqrtf(d2) < pl_cut (pp_thr): sqrtf is monotone, so
    // the lists' predicate sqrtf(d2) < cut is exactly d2 <= pl_d2 (pp_d2)
    float pl_d2, pp_d2;
    float hOverZeta, tempNorm;
    float h, zeta;
    // cadence
    int stride, nav, pairs_freq, pp_freq;
    long long end_step;               // the job stops after this many steps
    // heating
    int tempFreq;
    float initialT, deltaT;
    // pulling (force-ramp)
    int pulledEnd, pullFreq;
    float k_trans, deltax;
    float pullVec[3], cant0[3];
    // indentation (continuous surface)
    float tipR, tipAp, tipBp, surfAp, surfBp;
    float indKs, tipZeta;
    float dir[3], chip0[3], micaR0[3], micaN[3];
    int moveSurface, fixTrans;
    long long retraction;
    // hydrodynamics
    float4* rforce;  // scaled noise of the step
    float4* mforce;  // forces (+ pulling) of the step
    float4* ci;      // TEA row sums of (D_ij/D_ii)^2, w: sum of D_ij/D_ii
    float* eps;
    float4* fdf;     // F + D F (exact path)
    float* chol_M;   // RPY matrix / its factor, row-major, chol_ld x chol_ld
    size_t chol_ld;
    float hi_a, hi_epsmax;
    int hi_eps_freq, hi_capricious, hi_chol_freq;
};

struct DevJobState {
    long long step;   // local step at the start of the current block
    float var;        // sqrt(2 T zeta / h)
    int overflow;     // bit0 possible list, bit1 Verlet list
    float cant[3];    // cantilever chip position (pulling)
    float extF[3];
    float tip[3], chip[3], micaR[3];
    float V, dx;
    float tipFav[4];  // last averaged tip force (x,y,z,|f|)
    float tipDr;      // |chip - tip| of the last tip update, before the tip moved
    float beta;       // TEA beta' (bdhitea.cu:162-170)
    double eps_traj;  // TEA epsilon of the last update
    float cell_org[3];  // possible-pairs cell grid of the last rebuild
    float cell_inv;     // 1 / cell edge
    int cell_dim[3];
};

// Jobs of this many beads and more build their possible-pairs list from a
// cell grid instead of the all-pairs scan (same list, same order).
constexpr int kCellMinBeads = 8192;
// Jobs of at most this many beads are stepped with their coordinates staged
// in LDS (16 bytes per bead: 80 KB, two workgroups per CU).
constexpr int kLdsStepBeads = 5120;
// Beads per LDS-step unit (one 1024-thread workgroup).
constexpr int kLdsUnit = 1024;
// Verlet rows are re-sorted by list length at every kSortEvery-th rebuild.
constexpr int kSortEvery = 20;

// List entry k of the Verlet list (slot-major) / possible list (bead-major).
__device__ __forceinline__ int list_at(const void* p, int idx16, size_t at) {
    return idx16 ? (int)static_cast<const unsigned short*>(p)[at] : static_cast<const int*>(p)[at];
}
__device__ __forceinline__ void list_put(void* p, int idx16, size_t at, int v) {
    if (idx16) static_cast<unsigned short*>(p)[at] = (unsigned short)v;
    else static_cast<int*>(p)[at] = v;
}

}  // namespace sop
