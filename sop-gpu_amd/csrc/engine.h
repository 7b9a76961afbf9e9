// engine.h -- the device runtime: an HBM arena, a table of jobs, and the
// block loop that advances them.
//
// One Engine serves both ways of running:
//   * plain   -- run_job(): one job alone, kernels launched one by one in
//                runGPU's order (src/gsop.cu:111-163), the baseline;
//   * batched -- run_blocks(): every admitted job advances together, one
//                step kernel per step for all of them, replayed from a
//                hipGraph.
// Both launch the same kernels on the same per-job arrays, so a job's
// trajectory does not depend on which way it ran or what ran beside it.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/sopgpu.h"
#include "hi.h"
#include "hostutil.h"
#include "jobdesc.h"
#include "kernels.h"

namespace sop {

struct Error {
    int code;
    std::string msg;
};

// First-fit sub-allocator over device chunks, bounded by a byte budget.
class Arena {
   public:
    explicit Arena(size_t budget) : budget_(budget) {}
    ~Arena();
    void* alloc(size_t bytes);  // nullptr when the budget is exhausted
    void release(void* p);
    // Returns the chunks that hold no live block to the device; the bytes
    // they reserved count against the budget no longer.  Returns the bytes
    // freed.
    size_t trim();
    size_t reserved() const { return reserved_; }
    size_t in_use() const { return in_use_; }

   private:
    struct Chunk {
        char* base;   // null once trimmed (the slot is kept: live_ holds chunk indices)
        size_t size;
        size_t live = 0;  // blocks allocated from the chunk
        std::map<size_t, size_t> free;  // offset -> length
    };
    std::vector<Chunk> chunks_;
    std::map<void*, std::pair<int, size_t>> live_;  // ptr -> (chunk, length)
    size_t budget_, reserved_ = 0, in_use_ = 0;
};

struct HostJob {
    int id = -1;
    int n = 0;
    DevJob dj{};
    DevJobState st0{};
    long long step = 0;      // host mirror of the device step counter
    long long end_step = 0;
    std::vector<void*> blocks;
    unsigned long long bytes = 0;
    std::vector<size_t> sizes;  // bytes of each block (snapshots for the autotuner)
    HiExactWork hw{};      // exact-HI factorisation workspace (hw.M null otherwise)
    int tile0 = 0, ntile = 0;  // the job's range in the tile table
};

// Launch shapes the autotuner searches; none of them changes a result.
struct Tuning {
    int max_blocks;    // persistent grid of the global-gather step kernel (0: one workgroup per 4 tiles)
    int unit_rows;     // Verlet rows per LDS step unit (threads of its workgroup)
    int shared_units;  // rows left over by LDS jobs packed into shared units (else one job per unit)
    int pair_batch;    // Verlet entries fetched ahead per lane in the LDS step
    int verlet_chunk;  // 64-entry chunks in flight per row in the LDS Verlet build
    int sort_every;    // Verlet rows re-sorted at every k-th Verlet rebuild
    int split_rows;    // 4, 8 or 16: LDS-step jobs stepped by k_step_split, that many rows per workgroup (0: off)
    int split_global;  // the same for the jobs above the LDS limit (gathers from global memory)
};

class Engine {
   public:
    Engine(int device, size_t budget);
    ~Engine();

    int add_job(const sop_params& p, const sop_system& s, long long numsteps, int* id);
    int remove_job(int id);
    // Frees arena chunks that hold no job's arrays (see Arena::trim).
    size_t trim_arena() { return arena_.trim(); }
    static int job_bytes(const sop_params& p, const sop_system& s, unsigned long long* bytes);

    int run_blocks(int nblocks);
    int run_job(int id);
    int synchronize();
    int block_steps() const { return L_; }

    HostJob* find(int id);
    int slot_of(int id) const;
    const std::vector<HostJob>& jobs() const { return jobs_; }

    // read-back / hooks
    int get_coords(int id, float* out4);
    int set_coords(int id, const float* in4);
    // coordinates of local step s (inside a block: the buffer step s reads)
    int get_coords_at(int id, long long s, float* out4);
    // forces added to every bead in the job's next step (host potentials)
    int upload_forces(int id, const float* f4);
    int get_array(int id, int what, void* out);
    int list_cap(int id, int which, int* cap);
    int get_list(int id, int which, int* counts, int* entries);
    int compute_forces(int id, float* out4);
    int compute_energies(int id, float* per_bead, float* sums3);
    int rebuild_lists(int id, int mask);
    int job_status(int id, long long* step, long long* end, int* overflow);
    int get_indentation(int id, float* out13);
    int list_stats(double* avg4, long long* beads);

    // Staged block (the reference's updater passes one at a time, over all
    // jobs, with the same device-side cadence checks as launch_block):
    int stage_job_update(int mask);
    int stage_possible();
    int stage_verlet();
    int stage_tea_update();
    int stage_step(int kstep);
    int stage_advance(int nsteps);
    int read_state_of(int id, DevJobState* out);

    void set_tuning(int max_blocks, int use_graph);
    // Possible-pairs cell grid: jobs of min_beads beads and more admitted
    // from now on get one (0 keeps the threshold, < 0 gives none); enabled = 0
    // builds every list by the all-pairs scan.
    void set_cells(int min_beads, int enabled);
    // Step jobs of at most kLdsStepBeads beads with their coordinates in
    // LDS (default) or, enabled = 0, all from global memory.
    void set_lds_step(int enabled);
    void set_exact_mode(int mode) { exact_mode_ = mode; }
    int get_hi(int id, double* eps, float* beta);
    int get_hi_factor(int id, float* L);
    // Searches the launch shapes on the admitted jobs: each candidate runs
    // real blocks (whole list periods) from a snapshot of every job's device
    // arrays, which is restored afterwards, so the jobs do not advance.
    int autotune(Tuning* best, double* ms_default, double* ms_best, int* trials);
    Tuning tuning() const;
    int set_shape(const Tuning& t);
    void profile(bool on);
    void profile_read(long long* launches, double* ms);

    hipStream_t stream() const { return stream_; }
    int device() const { return device_; }

   private:
    int sync_tables();
    int launch_block(int slot_only);  // slot_only < 0: all jobs
    int lists_due(int slot_only) const;
    std::vector<long long> block_key() const;
    int capture_graph(const std::vector<long long>& key);
    void drop_graph();
    int check(hipError_t e, const char* what);
    int read_state(int slot, DevJobState* out);
    int pull_states();
    int synced_jobs_ = 0;

    int device_;
    hipStream_t stream_ = nullptr;
    hipStream_t aux_stream_ = nullptr;  // trailing-update stream of the exact-HI factorisation
    hipEvent_t ev_f_ = nullptr, ev_t2_[2] = {nullptr, nullptr};
    Arena arena_;
    std::vector<HostJob> jobs_;
    int next_id_ = 1;
    // device tables
    DevJob* d_jobs_ = nullptr;
    DevJobState* d_state_ = nullptr;
    int2* d_tiles_ = nullptr;
    int cap_jobs_ = 0, cap_tiles_ = 0, ntiles_ = 0;
    bool tables_dirty_ = true;
    int L_ = 1;
    // single-job tables for plain runs
    int2* d_tiles1_ = nullptr;
    int cap_tiles1_ = 0;
    int tiles1_n_ = -1;  // bead count the single-job tile list was built for
    int tiles_for_slot(int slot, int* nt);
    int plan_lds_units(std::vector<int2>& units, std::vector<int4>& segs) const;
    // tuning
    int max_blocks_ = 0;
    int cell_min_ = kCellMinBeads;
    bool use_cells_ = true;
    int cell_words_ = 0;
    int any_idx32_ = 0;  // some job stores int32 list entries
    bool use_lds_step_ = true;
    int2* d_units_ = nullptr;  // row-sort units (job slot, first row) per kLdsUnit rows of LDS-step jobs
    int cap_units_ = 0, nunits_ = 0, lds_beads_ = 0, any_global_ = 1;
    int2* d_tea_tiles_ = nullptr;  // tiles of the TEA jobs
    int cap_tea_tiles_ = 0, ntea_tiles_ = 0;
    int2* d_units1_ = nullptr;  // row-sort units of the single-job path
    int cap_units1_ = 0;
    // LDS step plan: units (first segment, segments) over segments (job
    // slot, first row, rows, LDS offset), for all jobs and for one job
    int2* d_sunits_ = nullptr;
    int4* d_segs_ = nullptr;
    int cap_sunits_ = 0, cap_segs_ = 0, nsunits_ = 0;
    int2* d_sunits1_ = nullptr;
    int4* d_segs1_ = nullptr;
    int cap_sunits1_ = 0, cap_segs1_ = 0, nsunits1_ = 0;
    int tiles1_rows_ = 0;  // unit rows the single-job plan was built for
    StepShape shape_;
    bool shared_units_ = true;
    // autotuner snapshot of the jobs' device arrays and state table
    int snapshot(void** buf);
    int restore(void* buf, const std::vector<long long>& steps);
    int time_blocks(int nblocks, double* ms);
    bool uses_cells(size_t n) const;
    int cell_words_of(const DevJob& J) const;
    int exact_mode_ = SOP_CHOL_BF16X3;
    bool any_tea_ = false, any_exact_ = false;
    bool use_graph_ = true;
    // captured blocks by launch sequence (block_key()): the list kernels due
    // and each exact-HI job's refactorisation steps; at most kMaxGraphs
    static constexpr size_t kMaxGraphs = 32;
    std::map<std::vector<long long>, hipGraphExec_t> graphs_;
    // profiling
    bool profiling_ = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pool_;
    std::vector<int> ev_launches_;  // step launches each event pair spans
    size_t ev_used_ = 0;
    long long prof_launches_ = 0;
    double prof_ms_ = 0.0;
    void flush_profile();
    float4* d_scratch4_ = nullptr;
    size_t scratch4_n_ = 0;
};

float sqrt_below_bound(float t);

}  // namespace sop
