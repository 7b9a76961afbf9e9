/*
 * sopgpu.h -- C ABI of the MI355X SOP Langevin engine.
 *
 * Plain pointers and sizes only; all device memory is owned by the engine.
 * Every function returns 0 on success and a negative SOP_E* code on
 * failure, with the message in sop_last_error().  The reference reports
 * the same conditions with DIE() and exit(-1) (src/Util/wrapper.h:23,
 * src/Util/wrapper.cpp:28-42); here they are per-call errors so a serving
 * loop can fail one job and keep the others.
 *
 * Reference interfaces each group replaces (paths relative to
 * BarsegovGroup/SOP-GPU):
 *   engine / jobs   GSOP state + initGPU/initFF/runGPU, src/gsop.h:39-67,
 *                   src/gsop.cu:45-163; Potential/Updater/Integrator
 *                   plugins src/gsop.h:69-105
 *   lists           CovalentPotential::buildMap covalent.cu:65-109,
 *                   NativePotential::buildMap native.cu:58-107,
 *                   PossiblepairList::update possiblepairlist.cu:41-48,
 *                   PairList::update pairlist.cu:100-128
 *   energies        SOPPotential::computeEnergy / sumEnergies gsop.cu:255-271
 *   RNG             initRand / generateSeeds / rforce, src/Integrators/ht.cu
 *   config runs     sop-gpu main(), src/sop-gpu.cpp:18-65
 *   plugins         SOPPotential / SOPUpdater / SOPIntegrator registries
 *                   and potentialByName / updaterByName, src/gsop.h:69-105,
 *                   src/gsop.cu:31-34,80-101,237-253 (C++ classes over
 *                   these entry points: sop_plugin.h)
 *
 * Threading: a handle (engine, server, simulation) is used by one host
 * thread at a time -- none of them is thread-safe, as the reference's global
 * state is not; use one handle per thread.  The error message of
 * sop_last_error() is kept per thread.  Handles on the same GPU may live in
 * different threads: each has its own stream and device memory.
 */
#ifndef SOPGPU_H
#define SOPGPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SOP_OK = 0,
    SOP_E_ARG = -1,      /* bad argument */
    SOP_E_HIP = -2,      /* HIP runtime error */
    SOP_E_MEM = -3,      /* HBM budget exhausted */
    SOP_E_LIMIT = -4,    /* a topology limit (max_covalent, max_native, ...) exceeded */
    SOP_E_OVERFLOW = -5, /* a pairlist outgrew its capacity on the device */
    SOP_E_CONFIG = -6,   /* parameter / topology / coordinate file error */
    SOP_E_STATE = -7,    /* call not valid in the job's current state */
};

/* Parameters of one system, named after the reference's PARAMETER()s
 * (Appendix of each plugin header: covalent.h:12-14, native.h:11-12,
 * pairs.h:12-15, pairlist.h:12-14, possiblepairlist.h:14-15,
 * langevin.h:13-23, pulling.h:15-26, indentation.h:12-37, gsop.h:35,
 * output_manager.h:19, dcd_manager.h:13).  Field order matches the
 * oracle's or_params. */
typedef struct sop_params {
    float kspring_cov, R_limit;
    float R_limit_bond;
    float pairs_cutoff, el, a;
    float pairlist_cutoff, pairs_threshold;
    int pairs_freq, possiblepairs_freq;
    int max_covalent, max_native, max_pairs, max_possiblePairs;
    float timestep, zeta, temperature;
    int minimization;
    int seed; /* seed + firstrun, as initRand receives it (langevin.cu:23) */
    int heating, tempFreq;
    float initialT, deltaT;
    int nav, outputtiming, dcdfreq;
    int pulling;
    float deltax, k_trans, fconst;
    int pullFreq;
    int pullDirectionVector; /* 0 endToEnd, 1 vector */
    float pullVector[3];
    int fixedEnd, pulledEnd;
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
    int hi_chol_freq;       /* refactorise every k steps (reference: 1) */
    int hi_faithful_blocks; /* apply L's 3x3 blocks as integrateCholesky_L reads them */
    /* host plugins (sop_plugin.h): the smallest frequency of updaters the
     * engine does not run (0: none), which bounds the stride as every
     * updater's frequency does (src/gsop.cu:116-122); host_forces != 0: a
     * host potential adds forces every step (sop_sim_add_forces) */
    int updater_freq_min;
    int host_forces;
} sop_params;

/* Topology and coordinates of one system in .top order
 * (src/IO/topio.cpp:36-80): bonds, native contacts (r0, eh) and the
 * optional [ pairs ] section that seeds the first possible-pairs list. */
typedef struct sop_system {
    int n;
    const float* coords4; /* n x (x,y,z,w) */
    int nbonds;
    const int* bond_i;
    const int* bond_j;
    const float* bond_r0;
    int nnative;
    const int* nat_i;
    const int* nat_j;
    const float* nat_r0;
    const float* nat_eh;
    int npairs;
    const int* pair_i;
    const int* pair_j;
    int nfixed;
    const int* fixed;
    int npulled;
    const int* pulled;
    /* optional initial Hybrid Taus state, n x 4 words (NULL: the seed table
     * of `seed`); lets replica k of a many-runs set take beads
     * [k*n, (k+1)*n) of one seed table as the reference does
     * (src/Integrators/langevin.cu:23-24) */
    const uint32_t* rng_state4;
} sop_system;

typedef struct sop_engine sop_engine;

const char* sop_last_error(void);
const char* sop_version(void);

/* Engine holding any number of jobs on one GPU.  hbm_budget bounds the
 * arena all job arrays live in (0: 90% of the device's free memory). */
int sop_engine_create(int device, unsigned long long hbm_budget, sop_engine** out);
void sop_engine_destroy(sop_engine* e);

/* Admit a job: builds the slot-major maps, seeds the RNG, uploads, and runs
 * the pre-loop Verlet build of runGPU (gsop.cu:126).  numsteps follows the
 * reference: the job runs numsteps/stride+1 blocks of stride steps. */
int sop_engine_add_job(sop_engine* e, const sop_params* p, const sop_system* s, long long numsteps, int* job_id);
int sop_engine_remove_job(sop_engine* e, int job_id);
/* bytes the job would take in the arena */
int sop_engine_job_bytes(const sop_params* p, const sop_system* s, unsigned long long* bytes);

/* Launch shape: max_blocks caps the persistent step grid (0: one
 * workgroup per 4 tiles); use_graph != 0 replays each block from a
 * captured hipGraph. */
int sop_engine_set_tuning(sop_engine* e, int max_blocks, int use_graph);
/* Possible-pairs builder (src/Updaters/possiblepairlist_kernel.cu:18-71
 * scans all same-trajectory pairs): jobs of min_beads beads and more admitted
 * after this call (default 8192; 0 keeps the threshold, < 0 none) get a cell
 * grid and build the identical list, in the identical order, from the 27
 * cells around each bead.  enabled = 0 makes every job use the all-pairs
 * scan again (the cell grids stay allocated). */
int sop_engine_set_cells(sop_engine* e, int min_beads, int enabled);
/* Step kernel variant (the reference's potentials + integrator,
 * src/gsop.cu:136-142): jobs of at most 5120 beads are stepped with their
 * coordinates staged in LDS (default, enabled = 1) or, enabled = 0, with
 * every neighbour gather from global memory.  Results are identical. */
int sop_engine_set_lds_step(sop_engine* e, int enabled);
/* Launch shapes of the step and list kernels; none changes a result (the
 * trajectories are the same bits under every shape).  No counterpart in the
 * reference, whose launches are one thread per bead in blocks of
 * `block_size` (src/gsop.cuh:13). */
typedef struct sop_tuning {
    int max_blocks;   /* persistent grid of the global-gather step kernel (jobs above 5120 beads; 0: one
                         workgroup per 4 tiles) */
    int unit_rows;    /* Verlet rows per LDS step unit, the threads of its workgroup: a multiple of 64, 64..1024 */
    int shared_units; /* rows left over by LDS jobs packed together into shared units (1) or one job per unit (0) */
    int pair_batch;   /* Verlet entries fetched ahead per lane in the LDS step: 4, 6 or 8 */
    int verlet_chunk; /* 64-entry chunks of a possible-list row in flight in the LDS Verlet build: 4 or 8 */
    int sort_every;   /* Verlet rows re-sorted by length at every k-th Verlet rebuild */
    int split_rows;   /* 4, 8 or 16: the jobs of at most 5120 beads are stepped with each bead's force terms
                         spread over a 1024-thread workgroup of that many beads (for a few small jobs; the
                         same results); 0: one bead per lane */
    int split_global; /* the same for the jobs above 5120 beads (partners gathered from global memory): 4, 8
                         or 16 beads per workgroup, 0: one bead per lane */
} sop_tuning;
int sop_engine_get_shape(sop_engine* e, sop_tuning* out);
int sop_engine_set_shape(sop_engine* e, const sop_tuning* shape);
/* Searches the shapes on the admitted jobs and keeps the fastest: each
 * candidate runs real blocks (whole periods of the jobs' list cadences)
 * from a device snapshot of every job, restored afterwards, so the jobs do
 * not advance.  Reports the chosen shape, the ms per step of the shape the
 * search started from and of the chosen one, and the candidates timed. */
int sop_engine_autotune(sop_engine* e, sop_tuning* best, double* ms_per_step_start, double* ms_per_step_best,
                        int* trials);

/* Batched stepping: every unfinished job advances one block of L steps
 * (L = gcd of the jobs' strides) per call, all jobs in one step kernel
 * per step. */
int sop_engine_run_blocks(sop_engine* e, int nblocks);
/* Plain stepping: run one job alone to its end, one launch per kernel,
 * in runGPU's order (the one-job-per-launch baseline). */
int sop_engine_run_job(sop_engine* e, int job_id);
int sop_engine_block_steps(sop_engine* e, int* L);
int sop_engine_synchronize(sop_engine* e);

int sop_engine_job_status(sop_engine* e, int job_id, long long* step, long long* end_step, int* overflow);
int sop_engine_active_jobs(sop_engine* e, int* count, int* max_n);

/* Read back (synchronous). */
int sop_engine_get_coords(sop_engine* e, int job_id, float* out4);
int sop_engine_set_coords(sop_engine* e, int job_id, const float* in4);
int sop_engine_get_temperature(sop_engine* e, int job_id, float* out);
int sop_engine_get_rng(sop_engine* e, int job_id, uint32_t* out4);
int sop_engine_get_ext(sop_engine* e, int job_id, float* out4);
/* which: 0 covalent, 1 native, 2 Verlet pairs, 3 possible pairs.
 * counts[n]; entries[cap*n] slot-major; *cap receives the capacity. */
int sop_engine_list_cap(sop_engine* e, int job_id, int which, int* cap);
int sop_engine_get_list(sop_engine* e, int job_id, int which, int* counts, int* entries);
/* tip(3), chip(3), micaR(3), tip force average (4) */
int sop_engine_get_indentation(sop_engine* e, int job_id, float* out13);

/* Parity hooks: forces from the current coordinates and lists (no state
 * change), per-bead energies (3n) and per-system sums (3) as
 * SOPPotential::sumEnergies forms them. */
int sop_engine_compute_forces(sop_engine* e, int job_id, float* out4);
int sop_engine_compute_energies(sop_engine* e, int job_id, float* per_bead3n, float* sums3);
/* Force a list rebuild now (mask: 1 possible, 2 Verlet). */
int sop_engine_rebuild_lists(sop_engine* e, int job_id, int mask);

/* Timing of the step kernel with HIP events on the engine's stream:
 * while enabled, each step-kernel launch is bracketed by events. */
int sop_engine_profile(sop_engine* e, int enable);
int sop_engine_profile_read(sop_engine* e, long long* launches, double* total_ms);
/* Average list lengths over admitted jobs' beads (covalent, native,
 * Verlet, possible) and the bead count, for the roofline byte count. */
int sop_engine_list_stats(sop_engine* e, double* avg4, long long* beads);

/* Hydrodynamics (src/Integrators/bdhitea.cu).  The exact path's trailing
 * update runs as split-BF16 three-product MFMA (mode 0, default) or FP32
 * MFMA (mode 1); the 64x64 diagonal blocks are factorised, and inverted for
 * the panel solves, in FP32 in mode 0 and in FP64 in mode 1. */
int sop_engine_set_chol_mode(sop_engine* e, int mode);
/* epsilon and beta' of the job's last TEA update (TeaUpdater::update,
 * bdhitea.cu:133-176) */
int sop_engine_get_hi(sop_engine* e, int job_id, double* epsilon, float* beta);
/* the job's current Cholesky factor, 3n x 3n row-major, upper part zero */
int sop_engine_get_hi_factor(sop_engine* e, int job_id, float* L);
/* Standalone factorisation of an n x n SPD matrix (row-major; the lower
 * triangle is read) into L (n x n, upper zero); with repeats > 0, one
 * untimed run and then `repeats` runs timed with HIP events, average in
 * *ms_per_factor.  Replaces the one-block-per-
 * system integrateCholesky_decompose (bdhitea_kernel.cu:154-190), which is
 * limited to 3N <= 384. */
int sop_chol_factor(int device, int n, const float* A, int mode, float* L, int repeats, double* ms_per_factor);
/* The same factorisation with nothing allocated.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * M is device memory the caller owns: an n x n SPD matrix, row-major with
 * leading dimension ld floats (the lower triangle is read), with room for mp
 * rows, mp = n rounded up to a multiple of 64, ld >= mp, ld a multiple of 4
 * and M 16-byte aligned.  On return (in
 * `stream` order; the call does not wait) its lower triangle holds L, equal
 * to sop_chol_factor's bit for bit; the upper triangle is left as it was and
 * rows n..mp-1 hold the identity padding.  stream is a hipStream_t (NULL:
 * the null stream); `device` becomes the calling thread's current device.
 * The workspace is one static device array per GPU (about 57 MB, held until
 * the process ends), enough for n <= 12288; a larger n is refused with
 * SOP_E_ARG.  Calls on one GPU share that array and the last call's contents
 * stay in it: calls that overlap in time on one GPU race, so issue them on
 * one stream or synchronise between them. */
int sop_chol_factor_noalloc(int device, int n, float* M, unsigned long long ld, int mode, void* stream);

/* Streaming copy of `bytes` (read once, written once) repeated `repeats`
 * times: achieved HBM GB/s (read + write), the reference point for the
 * roofline and for scaling memory counters. */
int sop_copy_bandwidth(int device, unsigned long long bytes, int repeats, double* gbps);

/* RNG (src/Integrators/ht.cu): seed table on the host with the reference's
 * ran2 stream, and a device probe that advances `count` draws per bead and
 * returns the uint32 outputs (gaussian=0) or Gaussian triples (gaussian=1). */
int sop_generate_seeds(int seed, int n, uint32_t* out4);
int sop_rng_probe(int device, const uint32_t* seeds4, int n, int count, int gaussian, float* out, uint32_t* state_out4);

/* Self-check with no counterpart in the reference: the repulsive-pair
 * kernel divides a^2 / d^2 by a short correctly rounded sequence (hardware
 * reciprocal estimate, one Newton step, Markstein's correction) whenever
 * d^2 lies in [2^-64, 2^62]; this counts, over every float bit pattern b in
 * [lo, hi) within that range and each numerator a[0..na) (each 0 or in
 * [2^-60, 2^60]), the quotients that differ from the IEEE division. */
int sop_check_division(int device, unsigned int lo, unsigned int hi, const float* a, int na,
                       unsigned long long* mismatches);

/* The sop-gpu program (src/sop-gpu.cpp:18-88, runGPU in src/gsop.cu:111-163):
 * reads a parameter file in the reference's format, its .top topology and
 * PDB/XYZ coordinates, runs the configured protocol (Langevin or BDHI,
 * pulling, indentation, heating, many runs per GPU) and writes the
 * reference's output files (energy .dat, pull/indentation .dat, DCD,
 * restart / reference / final PDBs) and screen log.  names/values override
 * parameters of the file.  Returns SOP_E_CONFIG with a message on a bad
 * parameter, topology or coordinate file. */
int sop_run_config(const char* path, int noverride, const char* const* names, const char* const* values);

/* The sop-gpu program with its plugin registries (src/gsop.cu:45-163).
 * sop_sim_create reads the parameter file (names/values override it), the
 * topology and coordinates, and builds the built-in potentials, updaters and
 * integrator in initFF's order (gsop.cu:80-101); the host may then append
 * its own plugins; sop_sim_run admits the trajectories on the GPU, runs
 * runGPU's loop and writes the outputs; it may be called once.  Errors:
 * SOP_E_CONFIG with the message (a bad parameter, a plugin's failure). */
typedef struct sop_sim sop_sim;
int sop_sim_create(const char* path, int noverride, const char* const* names, const char* const* values,
                   sop_sim** out);
int sop_sim_run(sop_sim* s);
void sop_sim_destroy(sop_sim* s);

/* Host plugins (SOPUpdater / SOPPotential, src/gsop.h:69-94), appended to
 * the registries in call order.  update() is called every stride steps and
 * checks its own frequency (sop_sim_state), compute() every step before
 * the integrator; a nonzero return ends the run with the message given to
 * sop_sim_fail().  energy(traj, id) gives the potential's term id of
 * trajectory traj for the energy .dat (energies_count terms, 0 for none).
 * release(user), if given, is called when the simulation is destroyed, or
 * at once when registration fails.  Registration is refused once running. */
typedef int (*sop_plugin_fn)(sop_sim* s, void* user);
typedef float (*sop_energy_fn)(sop_sim* s, int traj, int id, void* user);
typedef void (*sop_release_fn)(void* user);
int sop_sim_add_updater(sop_sim* s, const char* name, int frequency, sop_plugin_fn update, void* user,
                        sop_release_fn release);
int sop_sim_add_potential(sop_sim* s, const char* name, int energies_count, sop_plugin_fn compute,
                          sop_energy_fn energy, void* user, sop_release_fn release);
/* A plugin's failure message (the DIE() of the reference). */
void sop_sim_fail(const char* message);

/* The registries: counts, names and frequencies in order, and the lookups
 * of potentialByName / updaterByName (SOP_E_ARG "Unable to find requested
 * potential 'x'" when absent); the integrator's name and time step. */
int sop_sim_potential_count(const sop_sim* s, int* count);
int sop_sim_updater_count(const sop_sim* s, int* count);
const char* sop_sim_potential_name(const sop_sim* s, int index);
const char* sop_sim_updater_name(const sop_sim* s, int index);
int sop_sim_updater_frequency(const sop_sim* s, int index, int* frequency);
int sop_sim_potential_by_name(const sop_sim* s, const char* name, int* index);
int sop_sim_updater_by_name(const sop_sim* s, const char* name, int* index);
const char* sop_sim_integrator_name(const sop_sim* s);
/* SOPPotential::getEnergy of potential `index` at the last output step. */
int sop_sim_get_energy(sop_sim* s, int index, int traj, int id, float* energy);

/* State for plugins: the current step, beads per trajectory and
 * trajectories; the coordinates of trajectory traj now (n x 4); forces to
 * add to trajectory traj in the current step (n x 4, w ignored; summed over
 * the potentials that add them, applied by the step's integrator). */
int sop_sim_state(const sop_sim* s, long long* step, int* n, int* ntr);
int sop_sim_get_coords(sop_sim* s, int traj, float* out4);
int sop_sim_add_forces(sop_sim* s, int traj, const float* f4);

/* Continuous-batching request loop (no counterpart in the reference, which
 * runs one system or identical replicas per process, src/sop-gpu.cpp:18-65).
 * submit copies the request and queues it; pump admits queued requests in
 * arrival order while the HBM budget has room, advances every admitted job
 * by nblocks blocks in one step kernel per step, and retires finished jobs
 * (coordinates copied out, memory released).  Times are seconds since the
 * server was created; state: 0 queued, 1 running, 2 done, -1 failed. */
typedef struct sop_server sop_server;
int sop_server_create(int device, unsigned long long hbm_budget, sop_server** out);
void sop_server_destroy(sop_server* s);
int sop_server_set_tuning(sop_server* s, int max_blocks, int use_graph);
/* The launch shapes of the server's engine (sop_engine_set_shape), and
 * the autotuner on the jobs admitted so far (sop_engine_autotune). */
int sop_server_set_shape(sop_server* s, const sop_tuning* shape);
int sop_server_autotune(sop_server* s, sop_tuning* best, double* ms_per_step_start, double* ms_per_step_best,
                        int* trials);
int sop_server_submit(sop_server* s, const sop_params* p, const sop_system* sys, long long numsteps, int* ticket);
int sop_server_pump(sop_server* s, int nblocks, int* running, int* queued);
int sop_server_poll(sop_server* s, int ticket, int* state, double* t_submit, double* t_start, double* t_done);
int sop_server_result(sop_server* s, int ticket, float* coords4);
int sop_server_stats(sop_server* s, double* bead_steps, double* busy_seconds, int* admitted, int* finished,
                     unsigned long long* peak_bytes);

/* Synthetic systems: a seeded chain of n beads (3.8 A bonds, compact
 * globule) with native contacts |i-j|>2 and r<8 A, eh in [eh_lo, eh_hi].
 * Arrays are allocated by the library; free with sop_free_system. */
typedef struct sop_synth {
    int n;
    float* coords4;
    int nbonds;
    int* bond_i;
    int* bond_j;
    float* bond_r0;
    int nnative;
    int* nat_i;
    int* nat_j;
    float* nat_r0;
    float* nat_eh;
    int npairs; /* [ pairs ] of a .top file (initial possible list); 0 for synthetic systems */
    int* pair_i;
    int* pair_j;
} sop_synth;
int sop_synth_chain(int n, unsigned long long seed, float density, float eh_lo, float eh_hi, sop_synth* out);
/* nchains chains of n beads each placed on a sphere of radius R,
 * inter-chain contacts added for r < 8 A. */
int sop_synth_shell(int nchains, int n, unsigned long long seed, float radius, float eh_lo, float eh_hi, sop_synth* out);
void sop_free_system(sop_synth* s);

/* File formats of the reference.
 * sop_read_topology: SOP::load (src/IO/topio.cpp:36-80) of a .top file and,
 *   when coords_path is not NULL, readCoord (src/IO/pdbio.cpp:292-337) of a
 *   PDB (ATOM records) or .xyz file into coords4 (w = 0); free with
 *   sop_free_system.
 * sop_write_pdb: savePDB (src/IO/pdbio.cpp:194-219) of the .top's atoms at
 *   coords4 (hexadecimal serials from 100000 atoms on, CONECT records of the
 *   bonds).
 * sop_write_dcd: a DCD file (src/IO/dcdio.cpp:100-194) with the header of
 *   DcdOutputManager (NFILE = NPRIV = NSAVC = freq, src/Updaters/dcd_manager.cpp:48-52)
 *   and nframes frames; frames holds, per frame, X[natoms], Y[natoms], Z[natoms]. */
int sop_read_topology(const char* top_path, const char* coords_path, sop_synth* out);
int sop_write_pdb(const char* top_path, const float* coords4, int n, const char* out_path);
int sop_write_dcd(const char* path, int natoms, int nframes, const float* frames, int freq, float delta);

#ifdef __cplusplus
}
#endif
#endif
