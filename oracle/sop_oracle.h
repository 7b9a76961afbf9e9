/*
 * sop_oracle.h -- CPU restatement of the SOP-GPU hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls
 * this code: it is the checker used by tests/, by __graft_entry__.smoke()
 * and by the cpu_baseline leg of bench.py.
 *
 * Each function restates the reference algorithm named in its comment
 * (paths are relative to the reference repository BarsegovGroup/SOP-GPU).
 * The FP32 functions keep the reference's expression order and are built
 * with -ffp-contract=off, so they are the bit-level statement of what the
 * kernels compute; the FP64 functions evaluate the same model in double and
 * are the accuracy oracle.
 *
 * Pinning (what checks each function, see tests/test_oracle_*.py):
 *   ran2 / seed tables      -> reference src/Integrators/ran2.cpp compiled
 *                              as-is into oracle/_ref (bit-exact), plus the
 *                              Numerical Recipes closed form for idum<=0.
 *   Hybrid Taus             -> independent closed form of each Tausworthe
 *                              component (L'Ecuyer 1996) and the LCG.
 *   log / sincos / Gaussian -> FP64 libm within a few ulp; moments and KS.
 *   forces                  -> closed forms (two-bead cases), F = -grad U by
 *                              finite differences of the FP64 energies,
 *                              Newton's third law, FP64 within 1e-5.
 *   pairlists               -> brute-force set definition written
 *                              independently in numpy (test side).
 *   integrator / cadence    -> Euler closed form on a harmonic dimer, the
 *                              reference step count numsteps+stride.
 *   indentation             -> closed-form wall forces for a single bead.
 *   cantilever, tip update, -> FP64 closed forms of the updater passes
 *   heating                    (test_oracle_pins.py): k (c_prev - x), the
 *                              tip's overdamped step, T0 + floor(s/f) dT.
 *   TEA step, D F, L g      -> FP64 numpy statements of the step from the
 *                              RPY matrix, LAPACK factor and pinned noise.
 *   pull / indentation rows -> the same closed forms (values printed).
 */
#ifndef SOP_ORACLE_H
#define SOP_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Parameters of one simulation, named after the reference's PARAMETERs. */
typedef struct or_params {
    /* src/Potentials/covalent.h:12-14 */
    float kspring_cov, R_limit;
    /* src/Potentials/native.h:11 */
    float R_limit_bond;
    /* src/Potentials/pairs.h:12-14 */
    float pairs_cutoff, el, a;
    /* src/Updaters/pairlist.h:12-13, possiblepairlist.h:14-15 */
    float pairlist_cutoff, pairs_threshold;
    int pairs_freq, possiblepairs_freq;
    /* list capacities */
    int max_covalent, max_native, max_pairs, max_possiblePairs;
    /* src/Integrators/langevin.h:13-23 */
    float timestep, zeta, temperature;
    int minimization;
    int seed; /* seed + firstrun, as passed to initRand */
    int heating, tempFreq;
    float initialT, deltaT;
    /* cadence: src/gsop.h:35, output_manager.h:19, dcd_manager.h:13 */
    int nav, outputtiming, dcdfreq;
    /* src/Potentials/pulling.h:15-26 */
    int pulling;
    float deltax, k_trans, fconst;
    int pullFreq;
    int pullDirectionVector; /* 0 = endToEnd, 1 = vector */
    float pullVector[3];
    int fixedEnd, pulledEnd;
    /* src/Potentials/indentation.h:12-37 (continuous surface only) */
    int indentation;
    float indChip[3], indTip[3], indDirection[3], indSurfR0[3], indSurfN[3];
    float indTipR, indKs, indDeltaX;
    int indMoveSurface, indFixTrans;
    float indTipSigma, indSurfSigma, indTipEl, indSurfEl;
    float indTipA, indTipB, indSurfA, indSurfB;
    float indTipZeta;
    int indOutputFreq;
    long long indRetractionStep;
    /* src/Integrators/bdhitea.h:15-21 */
    int hi_on, hi_exact;
    float hi_a;
    int hi_epsilon_freq, hi_capricious, hi_unlisted;
    float hi_epsmax;
    int hi_chol_freq;       /* Cholesky refactorisation interval (reference: every step) */
    int hi_faithful_blocks; /* apply L's 3x3 blocks as integrateCholesky_L reads them */
    /* smallest frequency of updaters added by the host (0: none): it bounds
     * the stride as every updater's frequency does; host_forces is not
     * modelled (the oracle has no host potentials) */
    int updater_freq_min;
    int host_forces;
} or_params;

typedef struct or_sim or_sim;

/* Topology arrays use the reference's .top order
 * (src/IO/topio.cpp:36-80).  ntr > 1 replicates the topology as the
 * reference's many-runs-per-GPU mode does (covalent.cu:85-96). */
or_sim* or_create(const or_params* p, int N, int ntr, const float* coords4,
                  int nb, const int* bi, const int* bj, const float* br0,
                  int nn, const int* ni, const int* nj, const float* nr0, const float* neh,
                  int np, const int* pi, const int* pj,
                  int nfixed, const int* fixed, int npulled, const int* pulled);
void or_destroy(or_sim* s);
const char* or_error(void);

int  or_n(const or_sim* s);
void or_get_coords(const or_sim* s, float* out4);
void or_set_coords(or_sim* s, const float* in4);
void or_get_T(const or_sim* s, float* out);
void or_get_rng(const or_sim* s, uint32_t* out4);
void or_get_ext(const or_sim* s, float* out4);
long long or_step(const or_sim* s);
/* which: 0 covalent, 1 native, 2 Verlet pairs, 3 possible pairs.
 * Writes counts[n] and the slot-major entries [k*n+i] for k<cap. */
int  or_get_list(const or_sim* s, int which, int* counts, int* entries);
int  or_list_cap(const or_sim* s, int which);
int  or_overflow(const or_sim* s);

/* One force evaluation from the current coordinates and lists, in the
 * reference's potential order (covalent, native, pairs, indentation). */
void or_forces_f32(or_sim* s, float* out4);
void or_forces_f64(or_sim* s, double* out3);
/* Per-bead energies for covalent / native / pairs (3*n floats). */
void or_energies_f32(or_sim* s, float* out3n);
void or_energies_f64(const or_sim* s, double* out3n);
/* Per-trajectory energies as SOPPotential::sumEnergies forms them. */
void or_sum_energies(or_sim* s, float* out3ntr);

void or_build_possible(or_sim* s);
void or_build_verlet(or_sim* s);

/* Runs runGPU (src/gsop.cu:111-163) for `numsteps`: numsteps+stride
 * steps in blocks of stride.  Returns the number of steps done. */
long long or_run(or_sim* s, long long numsteps);
/* FP64 twin of or_run (no noise: temperature must be 0 or minimization on). */
long long or_run_f64(or_sim* s, long long numsteps, double* coords_out3);
int or_stride(const or_sim* s);

/* HI pieces: per-trajectory epsilon and beta' of the last TEA update, and
 * the factor (3N x 3N row-major lower triangle, trajectory 0). */
void or_hi_tea(const or_sim* s, double* eps, float* beta);
int or_hi_factor(const or_sim* s, float* L);
/* Standalone RPY (D/D_ii) matrix of coordinates, FP64, 3n x 3n row-major. */
void or_rpy_matrix(int n, const float* coords4, double a, double* D);

/* Indentation tip state (x,y,z of tip, chip, mica point) */
void or_get_indentation(const or_sim* s, float* out9);

/* DAT output rows recorded at outputtiming: step, T, Ecov, Enat, Erep, Q.
 * Returns the number of rows (at most cap). */
int or_get_dat(const or_sim* s, double* out6, int cap);

/* Rows of the pull .dat (pulling.cu:166-187) and indentation .dat
 * (indentation.cu:497-536) the reference writes: per row traj, step,
 * |end-to-end|, end-to-end . n, F . n, F xyz (8 doubles); step, dx,
 * averaged |f|, |averaged f|, Ks |chip - tip|, its average, averaged f
 * xyz, tip xyz, chip xyz (15 doubles).  Return the row count. */
int or_get_pull_rows(const or_sim* s, double* out8, int cap);
int or_get_ind_rows(const or_sim* s, double* out15, int cap);

/* RNG pieces (src/Integrators/ht.cu, ran2.cpp) */
typedef struct { int idum2, iy, iv[32]; } or_ran2_state;
void or_ran2_init(or_ran2_state* st);
double or_ran2(int* idum, or_ran2_state* st);
void or_generate_seeds(int seed, int n, uint32_t* out4);
uint32_t or_hybrid_taus(uint32_t* s4);
float or_uint_to_float(uint32_t u);
float or_logf(float x);
void or_sincosf(float x, float* s, float* c);
void or_rforce(uint32_t* s4, float* out4);

/* Plain kernels on raw arrays (CPU baseline): one step of the fused
 * per-bead force + Langevin update over slot-major lists, OpenMP over
 * beads when nthreads > 1.  Returns wall seconds for `steps` steps. */
double or_bench_steps(or_sim* s, int steps, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
