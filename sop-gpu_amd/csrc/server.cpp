// server.cpp -- continuous-batching request loop over one Engine.
//
// Requests (a system, its parameters and a step count) queue in arrival
// order.  Every pump admits queued requests while the HBM budget has room,
// advances all admitted jobs together (one step kernel per step, graph
// replay), and retires finished jobs: their coordinates are copied out,
// their arena blocks freed, and the next requests can take the space.
// Requests that arrive while others run join at the next block boundary.
// The reference has no counterpart; it runs one system (or identical
// replicas) per process (src/sop-gpu.cpp:18-65).
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <vector>

#include "../../include/sopgpu.h"
#include "engine.h"

namespace {

struct Request {
    int ticket = 0;
    sop_params p{};
    long long numsteps = 0;
    int n = 0;
    std::vector<float> coords;
    std::vector<int> bi, bj, ni, nj, pi, pj, fixed, pulled;
    std::vector<float> br, nr, ne;
    std::vector<uint32_t> rng;
    unsigned long long bytes = 0;
    int state = 0;  // 0 queued, 1 running, 2 done, -1 failed
    int job = -1;
    double t_submit = 0, t_start = 0, t_done = 0;
    std::vector<float> result;
    std::string error;
    long long bead_steps = 0;

    sop_system view() const {
        sop_system s{};
        s.n = n;
        s.coords4 = coords.data();
        s.nbonds = (int)br.size();
        s.bond_i = bi.data();
        s.bond_j = bj.data();
        s.bond_r0 = br.data();
        s.nnative = (int)nr.size();
        s.nat_i = ni.data();
        s.nat_j = nj.data();
        s.nat_r0 = nr.data();
        s.nat_eh = ne.data();
        s.npairs = (int)pi.size();
        s.pair_i = pi.data();
        s.pair_j = pj.data();
        s.nfixed = (int)fixed.size();
        s.fixed = fixed.data();
        s.npulled = (int)pulled.size();
        s.pulled = pulled.data();
        s.rng_state4 = rng.empty() ? nullptr : rng.data();
        return s;
    }
};

template <typename T>
std::vector<T> copy_of(const T* p, int n) {
    return (p && n > 0) ? std::vector<T>(p, p + n) : std::vector<T>();
}

}  // namespace

struct sop_server {
    sop::Engine* eng = nullptr;
    unsigned long long budget = 0;
    std::map<int, Request> reqs;
    std::deque<int> queue;
    std::vector<int> running;
    int next_ticket = 1;
    std::chrono::steady_clock::time_point t0;
    double bead_steps = 0.0, busy = 0.0;
    int admitted = 0, finished = 0;
    unsigned long long in_use = 0, peak = 0;
    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

extern "C" {

int sop_server_create(int device, unsigned long long hbm_budget, sop_server** out) {
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
    auto* s = new sop_server;
    try {
        s->eng = new sop::Engine(device, (size_t)hbm_budget);
    } catch (const sop::Error& err) {
        delete s;
        return sop::fail(err.code, err.msg);
    }
    s->budget = hbm_budget;
    s->t0 = std::chrono::steady_clock::now();
    *out = s;
    return 0;
}

void sop_server_destroy(sop_server* s) {
    if (!s) return;
    delete s->eng;
    delete s;
}

int sop_server_submit(sop_server* s, const sop_params* p, const sop_system* sys, long long numsteps, int* ticket) {
    if (!s || !p || !sys || !ticket) return sop::fail(SOP_E_ARG, "null argument");
    Request r;
    r.ticket = s->next_ticket++;
    r.p = *p;
    r.numsteps = numsteps;
    r.n = sys->n;
    r.coords = copy_of(sys->coords4, 4 * sys->n);
    r.bi = copy_of(sys->bond_i, sys->nbonds);
    r.bj = copy_of(sys->bond_j, sys->nbonds);
    r.br = copy_of(sys->bond_r0, sys->nbonds);
    r.ni = copy_of(sys->nat_i, sys->nnative);
    r.nj = copy_of(sys->nat_j, sys->nnative);
    r.nr = copy_of(sys->nat_r0, sys->nnative);
    r.ne = copy_of(sys->nat_eh, sys->nnative);
    r.pi = copy_of(sys->pair_i, sys->npairs);
    r.pj = copy_of(sys->pair_j, sys->npairs);
    r.fixed = copy_of(sys->fixed, sys->nfixed);
    r.pulled = copy_of(sys->pulled, sys->npulled);
    if (sys->rng_state4) r.rng = copy_of(sys->rng_state4, 4 * sys->n);
    r.t_submit = s->now();
    const sop_system v = r.view();
    if (int rc = sop::Engine::job_bytes(r.p, v, &r.bytes)) return rc;
    if (r.bytes > s->budget) return sop::fail(SOP_E_MEM, "job larger than the server's HBM budget");
    *ticket = r.ticket;
    s->queue.push_back(r.ticket);
    s->reqs.emplace(r.ticket, std::move(r));
    return 0;
}

// One pump: admit, advance nblocks blocks, retire.
int sop_server_pump(sop_server* s, int nblocks, int* running, int* queued) {
    if (!s) return sop::fail(SOP_E_ARG, "null server");
    sop::Engine& E = *s->eng;
    // admission in arrival order while the budget has room
    while (!s->queue.empty()) {
        Request& r = s->reqs[s->queue.front()];
        if (s->in_use + r.bytes > s->budget) break;
        const sop_system v = r.view();
        int id = 0;
        int rc = E.add_job(r.p, v, r.numsteps, &id);
        if (rc == SOP_E_MEM) {
            // the budget has room by the byte count, but the arena's chunks
            // do not: return the empty ones to the device and try once more
            if (E.trim_arena() > 0) rc = E.add_job(r.p, v, r.numsteps, &id);
        }
        s->queue.pop_front();
        if (rc == SOP_E_MEM && !s->running.empty()) {  // retry after retirements free more chunks
            s->queue.push_front(r.ticket);
            break;
        }
        if (rc) {  // a bad request, or no room with nothing else admitted: it fails
            r.state = -1;
            r.error = sop_last_error();
            r.t_done = s->now();
            continue;
        }
        r.job = id;
        r.state = 1;
        r.t_start = s->now();
        s->in_use += r.bytes;
        s->peak = s->in_use > s->peak ? s->in_use : s->peak;
        s->running.push_back(r.ticket);
        s->admitted++;
    }
    if (!s->running.empty() && nblocks > 0) {
        const double a = s->now();
        const int L = E.block_steps();
        std::vector<long long> before;
        for (int t : s->running) before.push_back(E.find(s->reqs[t].job)->step);
        if (int rc = E.run_blocks(nblocks)) return rc;
        for (size_t k = 0; k < s->running.size(); ++k) {
            Request& r = s->reqs[s->running[k]];
            const long long done = E.find(r.job)->step - before[k];
            r.bead_steps += done * (long long)r.n;
            s->bead_steps += (double)done * r.n;
        }
        (void)L;
        // retire finished jobs (the read-back waits for the device)
        std::vector<int> still;
        for (int t : s->running) {
            Request& r = s->reqs[t];
            sop::HostJob* hj = E.find(r.job);
            if (hj && hj->step >= hj->end_step) {
                r.result.resize(4 * (size_t)r.n);
                int rc = E.get_coords(r.job, r.result.data());
                long long st = 0, end = 0;
                int ovf = 0;
                if (!rc) rc = E.job_status(r.job, &st, &end, &ovf);
                r.t_done = s->now();
                if (rc || ovf) {
                    r.state = -1;
                    r.error = rc ? sop_last_error() : "pairlist overflow or HI tensor check failed on the device";
                } else {
                    r.state = 2;
                }
                E.remove_job(r.job);
                s->in_use -= r.bytes;
                s->finished++;
            } else {
                still.push_back(t);
            }
        }
        s->running.swap(still);
        s->busy += s->now() - a;
    }
    if (running) *running = (int)s->running.size();
    if (queued) *queued = (int)s->queue.size();
    return 0;
}

int sop_server_poll(sop_server* s, int ticket, int* state, double* t_submit, double* t_start, double* t_done) {
    if (!s) return sop::fail(SOP_E_ARG, "null server");
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return sop::fail(SOP_E_ARG, "unknown ticket");
    const Request& r = it->second;
    if (state) *state = r.state;
    if (t_submit) *t_submit = r.t_submit;
    if (t_start) *t_start = r.t_start;
    if (t_done) *t_done = r.t_done;
    return 0;
}

int sop_server_result(sop_server* s, int ticket, float* coords4) {
    if (!s || !coords4) return sop::fail(SOP_E_ARG, "null argument");
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return sop::fail(SOP_E_ARG, "unknown ticket");
    const Request& r = it->second;
    if (r.state == -1) return sop::fail(SOP_E_STATE, "job failed: " + r.error);
    if (r.state != 2) return sop::fail(SOP_E_STATE, "job not finished");
    std::memcpy(coords4, r.result.data(), r.result.size() * sizeof(float));
    return 0;
}

int sop_server_stats(sop_server* s, double* bead_steps, double* busy_seconds, int* admitted, int* finished,
                     unsigned long long* peak_bytes) {
    if (!s) return sop::fail(SOP_E_ARG, "null server");
    if (bead_steps) *bead_steps = s->bead_steps;
    if (busy_seconds) *busy_seconds = s->busy;
    if (admitted) *admitted = s->admitted;
    if (finished) *finished = s->finished;
    if (peak_bytes) *peak_bytes = s->peak;
    return 0;
}

int sop_server_set_shape(sop_server* s, const sop_tuning* t) {
    if (!s || !t) return sop::fail(SOP_E_ARG, "null argument");
    return s->eng->set_shape(sop::Tuning{t->max_blocks, t->unit_rows, t->shared_units, t->pair_batch,
                                         t->verlet_chunk, t->sort_every, t->split_rows, t->split_global});
}

int sop_server_autotune(sop_server* s, sop_tuning* best, double* ms_start, double* ms_best, int* trials) {
    if (!s) return sop::fail(SOP_E_ARG, "null server");
    sop::Tuning t{};
    double a = 0.0, b = 0.0;
    int n = 0;
    if (int rc = s->eng->autotune(&t, &a, &b, &n)) return rc;
    if (best)
        *best = sop_tuning{t.max_blocks,   t.unit_rows,  t.shared_units, t.pair_batch,
                           t.verlet_chunk, t.sort_every, t.split_rows,   t.split_global};
    if (ms_start) *ms_start = a;
    if (ms_best) *ms_best = b;
    if (trials) *trials = n;
    return 0;
}

int sop_server_set_tuning(sop_server* s, int max_blocks, int use_graph) {
    if (!s) return sop::fail(SOP_E_ARG, "null server");
    s->eng->set_tuning(max_blocks, use_graph);
    return 0;
}

}  // extern "C"
