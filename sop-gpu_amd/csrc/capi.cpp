// capi.cpp -- extern "C" entry points (include/sopgpu.h).
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>
#include <vector>

#include "../../include/sopgpu.h"
#include "engine.h"
#include "hostutil.h"
#include "hi.h"
#include "kernels.h"

namespace sop {
const char* last_error();
}

struct sop_engine {
    sop::Engine* eng;
};

#define ENG(e)                                                   \
    if (!(e) || !(e)->eng) return sop::fail(SOP_E_ARG, "null engine"); \
    sop::Engine& E = *(e)->eng

extern "C" {

const char* sop_last_error(void) { return sop::last_error(); }
const char* sop_version(void) { return "sop-mi355x 0.1 (gfx950)"; }

int sop_engine_create(int device, unsigned long long hbm_budget, sop_engine** out) {
    if (!out) return sop::fail(SOP_E_ARG, "null output");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return sop::fail(SOP_E_HIP, "no HIP device available");
    if (device < 0 || device >= ndev) return sop::fail(SOP_E_ARG, "device index out of range");
    if (hbm_budget == 0) {
        size_t fr = 0, tot = 0;
        if (hipSetDevice(device) != hipSuccess || hipMemGetInfo(&fr, &tot) != hipSuccess)
            return sop::fail(SOP_E_HIP, "hipMemGetInfo failed");
        hbm_budget = (unsigned long long)(0.9 * (double)fr);
    }
    try {
        auto* e = new sop_engine;
        e->eng = new sop::Engine(device, (size_t)hbm_budget);
        *out = e;
        return 0;
    } catch (const sop::Error& err) {
        return sop::fail(err.code, err.msg);
    } catch (const std::bad_alloc&) {
        return sop::fail(SOP_E_MEM, "host allocation failed");
    }
}

void sop_engine_destroy(sop_engine* e) {
    if (!e) return;
    delete e->eng;
    delete e;
}

int sop_engine_add_job(sop_engine* e, const sop_params* p, const sop_system* s, long long numsteps, int* job_id) {
    ENG(e);
    if (!p || !s || !job_id) return sop::fail(SOP_E_ARG, "null argument");
    try {
        return E.add_job(*p, *s, numsteps, job_id);
    } catch (const std::bad_alloc&) {
        return sop::fail(SOP_E_MEM, "host allocation failed");
    }
}

int sop_engine_remove_job(sop_engine* e, int job_id) {
    ENG(e);
    return E.remove_job(job_id);
}

int sop_engine_job_bytes(const sop_params* p, const sop_system* s, unsigned long long* bytes) {
    if (!p || !s || !bytes) return sop::fail(SOP_E_ARG, "null argument");
    return sop::Engine::job_bytes(*p, *s, bytes);
}

int sop_engine_set_tuning(sop_engine* e, int max_blocks, int use_graph) {
    ENG(e);
    if (max_blocks < 0) return sop::fail(SOP_E_ARG, "max_blocks < 0");
    E.set_tuning(max_blocks, use_graph);
    return 0;
}

int sop_engine_set_lds_step(sop_engine* e, int enabled) {
    ENG(e);
    E.set_lds_step(enabled);
    return 0;
}

int sop_engine_set_cells(sop_engine* e, int min_beads, int enabled) {
    ENG(e);
    E.set_cells(min_beads, enabled);
    return 0;
}

static sop_tuning to_c(const sop::Tuning& t) {
    return sop_tuning{t.max_blocks, t.unit_rows, t.shared_units, t.pair_batch, t.verlet_chunk, t.sort_every,
                      t.split_rows, t.split_global};
}

int sop_engine_get_shape(sop_engine* e, sop_tuning* out) {
    ENG(e);
    if (!out) return sop::fail(SOP_E_ARG, "null output");
    *out = to_c(E.tuning());
    return 0;
}

int sop_engine_set_shape(sop_engine* e, const sop_tuning* t) {
    ENG(e);
    if (!t) return sop::fail(SOP_E_ARG, "null shape");
    return E.set_shape(sop::Tuning{t->max_blocks, t->unit_rows, t->shared_units, t->pair_batch, t->verlet_chunk,
                                   t->sort_every, t->split_rows, t->split_global});
}

int sop_engine_autotune(sop_engine* e, sop_tuning* best, double* ms_start, double* ms_best, int* trials) {
    ENG(e);
    sop::Tuning t{};
    double a = 0.0, b = 0.0;
    int n = 0;
    const int rc = E.autotune(&t, &a, &b, &n);
    if (rc) return rc;
    if (best) *best = to_c(t);
    if (ms_start) *ms_start = a;
    if (ms_best) *ms_best = b;
    if (trials) *trials = n;
    return 0;
}

int sop_engine_run_blocks(sop_engine* e, int nblocks) {
    ENG(e);
    if (nblocks < 0) return sop::fail(SOP_E_ARG, "nblocks < 0");
    return E.run_blocks(nblocks);
}

int sop_engine_run_job(sop_engine* e, int job_id) {
    ENG(e);
    return E.run_job(job_id);
}

int sop_engine_block_steps(sop_engine* e, int* L) {
    ENG(e);
    if (!L) return sop::fail(SOP_E_ARG, "null argument");
    *L = E.block_steps();
    return 0;
}

int sop_engine_synchronize(sop_engine* e) {
    ENG(e);
    return E.synchronize();
}

int sop_engine_job_status(sop_engine* e, int job_id, long long* step, long long* end_step, int* overflow) {
    ENG(e);
    return E.job_status(job_id, step, end_step, overflow);
}

int sop_engine_active_jobs(sop_engine* e, int* count, int* max_n) {
    ENG(e);
    int c = 0, m = 0;
    for (auto& j : E.jobs())
        if (j.step < j.end_step) {
            ++c;
            if (j.n > m) m = j.n;
        }
    if (count) *count = c;
    if (max_n) *max_n = m;
    return 0;
}

int sop_engine_get_coords(sop_engine* e, int job_id, float* out4) {
    ENG(e);
    if (!out4) return sop::fail(SOP_E_ARG, "null output");
    return E.get_coords(job_id, out4);
}

int sop_engine_set_coords(sop_engine* e, int job_id, const float* in4) {
    ENG(e);
    if (!in4) return sop::fail(SOP_E_ARG, "null input");
    return E.set_coords(job_id, in4);
}

int sop_engine_get_temperature(sop_engine* e, int job_id, float* out) {
    ENG(e);
    return E.get_array(job_id, 0, out);
}

int sop_engine_get_rng(sop_engine* e, int job_id, uint32_t* out4) {
    ENG(e);
    return E.get_array(job_id, 1, out4);
}

int sop_engine_get_ext(sop_engine* e, int job_id, float* out4) {
    ENG(e);
    return E.get_array(job_id, 2, out4);
}

int sop_engine_list_cap(sop_engine* e, int job_id, int which, int* cap) {
    ENG(e);
    if (!cap) return sop::fail(SOP_E_ARG, "null output");
    return E.list_cap(job_id, which, cap);
}

int sop_engine_get_list(sop_engine* e, int job_id, int which, int* counts, int* entries) {
    ENG(e);
    if (!counts || !entries) return sop::fail(SOP_E_ARG, "null output");
    return E.get_list(job_id, which, counts, entries);
}

int sop_engine_get_indentation(sop_engine* e, int job_id, float* out13) {
    ENG(e);
    if (!out13) return sop::fail(SOP_E_ARG, "null output");
    return E.get_indentation(job_id, out13);
}

int sop_engine_compute_forces(sop_engine* e, int job_id, float* out4) {
    ENG(e);
    if (!out4) return sop::fail(SOP_E_ARG, "null output");
    return E.compute_forces(job_id, out4);
}

int sop_engine_compute_energies(sop_engine* e, int job_id, float* per_bead3n, float* sums3) {
    ENG(e);
    return E.compute_energies(job_id, per_bead3n, sums3);
}

int sop_engine_rebuild_lists(sop_engine* e, int job_id, int mask) {
    ENG(e);
    return E.rebuild_lists(job_id, mask);
}

int sop_engine_profile(sop_engine* e, int enable) {
    ENG(e);
    E.profile(enable != 0);
    return 0;
}

int sop_engine_profile_read(sop_engine* e, long long* launches, double* total_ms) {
    ENG(e);
    long long l = 0;
    double ms = 0.0;
    E.profile_read(&l, &ms);
    if (launches) *launches = l;
    if (total_ms) *total_ms = ms;
    return 0;
}

int sop_engine_list_stats(sop_engine* e, double* avg4, long long* beads) {
    ENG(e);
    if (!avg4 || !beads) return sop::fail(SOP_E_ARG, "null output");
    return E.list_stats(avg4, beads);
}

int sop_engine_set_chol_mode(sop_engine* e, int mode) {
    ENG(e);
    if (mode != sop::SOP_CHOL_BF16X3 && mode != sop::SOP_CHOL_F32) return sop::fail(SOP_E_ARG, "unknown mode");
    E.set_exact_mode(mode);
    return 0;
}

int sop_engine_get_hi(sop_engine* e, int job_id, double* epsilon, float* beta) {
    ENG(e);
    return E.get_hi(job_id, epsilon, beta);
}

int sop_engine_get_hi_factor(sop_engine* e, int job_id, float* L) {
    ENG(e);
    if (!L) return sop::fail(SOP_E_ARG, "null output");
    return E.get_hi_factor(job_id, L);
}

int sop_chol_factor(int device, int n, const float* A, int mode, float* L, int repeats, double* ms_per_factor) {
    if (n <= 0 || !A || repeats < 0) return sop::fail(SOP_E_ARG, "bad arguments");
    if (mode != sop::SOP_CHOL_BF16X3 && mode != sop::SOP_CHOL_F32) return sop::fail(SOP_E_ARG, "unknown mode");
    if (hipSetDevice(device) != hipSuccess) return sop::fail(SOP_E_HIP, "hipSetDevice failed");
    const size_t mp = ((size_t)n + 63) / 64 * 64;
    std::vector<float> h(mp * mp, 0.0f);
    for (size_t r = 0; r < (size_t)n; ++r)
        for (size_t c = 0; c <= r; ++c) h[r * mp + c] = A[r * n + c];
    for (size_t r = n; r < mp; ++r) h[r * mp + r] = 1.0f;
    float *dA = nullptr, *dM = nullptr;
    sop::CholWork w{};
    hipStream_t s = nullptr, aux = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr, ef = nullptr, et0 = nullptr, et1 = nullptr;
    int rc = 0;
    hipError_t err = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (err == hipSuccess) err = hipStreamCreateWithFlags(&aux, hipStreamNonBlocking);
    if (err == hipSuccess) err = hipEventCreateWithFlags(&ef, hipEventDisableTiming);
    if (err == hipSuccess) err = hipEventCreateWithFlags(&et0, hipEventDisableTiming);
    if (err == hipSuccess) err = hipEventCreateWithFlags(&et1, hipEventDisableTiming);
    if (err == hipSuccess) err = hipMalloc(&dA, mp * mp * 4);
    if (err == hipSuccess) err = hipMalloc(&dM, mp * mp * 4);
    if (err == hipSuccess) err = hipMalloc(&w.linv, sop::kCholLinvBytes);
    if (err == hipSuccess) err = hipMalloc(&w.bbuf, sop::chol_bbuf_bytes(mp));
    if (err == hipSuccess) err = hipMalloc(&w.phi, sop::chol_panel_bytes(mp));
    if (err == hipSuccess) err = hipMalloc(&w.plo, sop::chol_panel_bytes(mp));
    if (err == hipSuccess) err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
    if (err == hipSuccess) err = hipMemcpy(dA, h.data(), mp * mp * 4, hipMemcpyHostToDevice);
    double total = 0.0;
    const int reps = repeats > 0 ? repeats : 1;
    // one untimed factorisation first when timing: the first launches of
    // the kernels load their code and would count in the average
    for (int r = repeats > 0 ? -1 : 0; r < reps && err == hipSuccess; ++r) {
        err = hipMemcpyAsync(dM, dA, mp * mp * 4, hipMemcpyDeviceToDevice, s);
        if (err == hipSuccess) err = hipEventRecord(e0, s);
        if (err == hipSuccess) err = sop::launch_chol_factor(dM, mp, (int)mp, mode, w, s, aux, ef, et0, et1);
        if (err == hipSuccess) err = hipEventRecord(e1, s);
        if (err == hipSuccess) err = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
        if (r >= 0) total += ms;
    }
    if (err == hipSuccess && L) {
        err = hipMemcpy(h.data(), dM, mp * mp * 4, hipMemcpyDeviceToHost);
        for (size_t r = 0; r < (size_t)n; ++r)
            for (size_t c = 0; c < (size_t)n; ++c) L[r * n + c] = c <= r ? h[r * mp + c] : 0.0f;
    }
    if (err != hipSuccess) rc = sop::fail(SOP_E_HIP, std::string("cholesky: ") + hipGetErrorString(err));
    if (ms_per_factor) *ms_per_factor = total / reps;
    for (hipEvent_t ev : {e0, e1, ef, et0, et1})
        if (ev) (void)hipEventDestroy(ev);
    if (aux) (void)hipStreamDestroy(aux);
    (void)hipFree(dA);
    (void)hipFree(dM);
    (void)hipFree(w.linv);
    (void)hipFree(w.bbuf);
    (void)hipFree(w.phi);
    (void)hipFree(w.plo);
    if (s) (void)hipStreamDestroy(s);
    return rc;
}

int sop_copy_bandwidth(int device, unsigned long long bytes, int repeats, double* gbps) {
    if (bytes < 16 || repeats <= 0 || !gbps) return sop::fail(SOP_E_ARG, "bad arguments");
    if (hipSetDevice(device) != hipSuccess) return sop::fail(SOP_E_HIP, "hipSetDevice failed");
    const size_t n = (size_t)(bytes / 16);
    float4 *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t err = hipMalloc(&a, n * 16);
    if (err == hipSuccess) err = hipMalloc(&b, n * 16);
    if (err == hipSuccess) err = hipMemset(a, 0, n * 16);
    if (err == hipSuccess) err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
    if (err == hipSuccess) err = sop::launch_copy(a, b, n, nullptr);  // warm-up
    if (err == hipSuccess) err = hipEventRecord(e0, nullptr);
    for (int r = 0; r < repeats && err == hipSuccess; ++r) err = sop::launch_copy(a, b, n, nullptr);
    if (err == hipSuccess) err = hipEventRecord(e1, nullptr);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
    int rc = 0;
    if (err != hipSuccess) rc = sop::fail(SOP_E_HIP, std::string("copy: ") + hipGetErrorString(err));
    else *gbps = 2.0 * (double)n * 16 * repeats / (ms * 1e-3) / 1e9;  // bytes read + written
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}

int sop_generate_seeds(int seed, int n, uint32_t* out4) {
    if (n < 0 || (n > 0 && !out4)) return sop::fail(SOP_E_ARG, "bad arguments");
    sop::generate_seeds(seed, n, out4);
    return 0;
}

int sop_rng_probe(int device, const uint32_t* seeds4, int n, int count, int gaussian, float* out, uint32_t* state_out4) {
    if (n <= 0 || count <= 0 || !seeds4 || !out) return sop::fail(SOP_E_ARG, "bad arguments");
    if (hipSetDevice(device) != hipSuccess) return sop::fail(SOP_E_HIP, "hipSetDevice failed");
    uint4* d_state = nullptr;
    float* d_out = nullptr;
    const size_t per = gaussian ? 3 : 1;
    const size_t outbytes = (size_t)n * count * per * 4;
    if (hipMalloc(&d_state, (size_t)n * 16) != hipSuccess || hipMalloc(&d_out, outbytes) != hipSuccess) {
        (void)hipFree(d_state);
        return sop::fail(SOP_E_HIP, "hipMalloc failed");
    }
    int rc = 0;
    hipError_t err = hipMemcpy(d_state, seeds4, (size_t)n * 16, hipMemcpyHostToDevice);
    if (err == hipSuccess) err = sop::launch_rng_probe(d_state, n, count, gaussian, d_out, nullptr);
    if (err == hipSuccess) err = hipDeviceSynchronize();
    if (err == hipSuccess) err = hipMemcpy(out, d_out, outbytes, hipMemcpyDeviceToHost);
    if (err == hipSuccess && state_out4) err = hipMemcpy(state_out4, d_state, (size_t)n * 16, hipMemcpyDeviceToHost);
    if (err != hipSuccess) rc = sop::fail(SOP_E_HIP, std::string("rng probe: ") + hipGetErrorString(err));
    (void)hipFree(d_state);
    (void)hipFree(d_out);
    return rc;
}

int sop_check_division(int device, unsigned int lo, unsigned int hi, const float* a, int na,
                       unsigned long long* mismatches) {
    if (!a || na <= 0 || !mismatches || hi < lo) return sop::fail(SOP_E_ARG, "bad arguments");
    // the range the step kernel sends to the short division
    const unsigned bmin = 0x1f800000u, bmax = 0x5e800000u;  // 2^-64, 2^62
    if (lo < bmin || hi > bmax + 1) return sop::fail(SOP_E_ARG, "b outside [2^-64, 2^62]");
    for (int k = 0; k < na; ++k)
        if (!(a[k] == 0.0f || (a[k] >= 0x1p-60f && a[k] <= 0x1p60f)))
            return sop::fail(SOP_E_ARG, "a outside {0} U [2^-60, 2^60]");
    if (hipSetDevice(device) != hipSuccess) return sop::fail(SOP_E_HIP, "hipSetDevice failed");
    float* d_a = nullptr;
    unsigned long long* d_bad = nullptr;
    if (hipMalloc(&d_a, (size_t)na * 4) != hipSuccess || hipMalloc(&d_bad, 8) != hipSuccess) {
        (void)hipFree(d_a);
        return sop::fail(SOP_E_HIP, "hipMalloc failed");
    }
    int rc = 0;
    hipError_t err = hipMemcpy(d_a, a, (size_t)na * 4, hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipMemset(d_bad, 0, 8);
    if (err == hipSuccess) err = sop::launch_division_check(lo, hi, d_a, na, d_bad, nullptr);
    if (err == hipSuccess) err = hipDeviceSynchronize();
    if (err == hipSuccess) err = hipMemcpy(mismatches, d_bad, 8, hipMemcpyDeviceToHost);
    if (err != hipSuccess) rc = sop::fail(SOP_E_HIP, std::string("division check: ") + hipGetErrorString(err));
    (void)hipFree(d_a);
    (void)hipFree(d_bad);
    return rc;
}

}  // extern "C"
