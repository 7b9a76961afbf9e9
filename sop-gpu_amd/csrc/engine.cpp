// engine.cpp -- job admission, HBM arena and the block loop.
#include "engine.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>

#include "hostutil.h"
#include "kernels.h"

namespace sop {

// Device copy of a host table, growing the buffer when needed.
template <typename T>
static hipError_t upload_grow(T*& d, int& cap, const std::vector<T>& v, hipStream_t stream) {
    const int m = (int)v.size();
    if (m > cap) {
        (void)hipFree(d);
        d = nullptr;
        cap = 0;
        const int c = std::max(64, 2 * m);
        if (hipError_t e = hipMalloc(&d, (size_t)c * sizeof(T))) return e;
        cap = c;
    }
    return m ? hipMemcpyAsync(d, v.data(), (size_t)m * sizeof(T), hipMemcpyHostToDevice, stream) : hipSuccess;
}


static thread_local std::string g_error;
static thread_local int g_code = 0;

void set_error(int code, const std::string& msg) {
    g_code = code;
    g_error = msg;
}
int fail(int code, const std::string& msg) {
    set_error(code, msg);
    return code;
}
const char* last_error() { return g_error.c_str(); }

// ---------------------------------------------------------------------------
// Arena
// ---------------------------------------------------------------------------
static constexpr size_t kAlign = 256;
static constexpr size_t kChunk = size_t(1) << 30;

Arena::~Arena() {
    for (auto& c : chunks_)
        if (c.base) (void)hipFree(c.base);
}

void* Arena::alloc(size_t bytes) {
    bytes = (bytes + kAlign - 1) / kAlign * kAlign;
    if (bytes == 0) bytes = kAlign;
    for (size_t ci = 0; ci < chunks_.size(); ++ci) {
        auto& c = chunks_[ci];
        if (!c.base) continue;  // trimmed
        for (auto it = c.free.begin(); it != c.free.end(); ++it) {
            if (it->second >= bytes) {
                size_t off = it->first, len = it->second;
                c.free.erase(it);
                if (len > bytes) c.free[off + bytes] = len - bytes;
                void* p = c.base + off;
                live_[p] = {(int)ci, bytes};
                c.live++;
                in_use_ += bytes;
                return p;
            }
        }
    }
    size_t want = std::max(bytes, kChunk);
    if (reserved_ + want > budget_) want = bytes;
    if (reserved_ + want > budget_) return nullptr;
    char* base = nullptr;
    if (hipMalloc(&base, want) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    Chunk c;
    c.base = base;
    c.size = want;
    c.live = 1;
    if (want > bytes) c.free[bytes] = want - bytes;
    chunks_.push_back(c);
    reserved_ += want;
    void* p = base;
    live_[p] = {(int)chunks_.size() - 1, bytes};
    in_use_ += bytes;
    return p;
}

void Arena::release(void* p) {
    auto it = live_.find(p);
    if (it == live_.end()) return;
    auto& c = chunks_[it->second.first];
    size_t off = (char*)p - c.base, len = it->second.second;
    in_use_ -= len;
    c.live--;
    live_.erase(it);
    auto nx = c.free.lower_bound(off);
    if (nx != c.free.end() && off + len == nx->first) {
        len += nx->second;
        nx = c.free.erase(nx);
    }
    if (nx != c.free.begin()) {
        auto pv = std::prev(nx);
        if (pv->first + pv->second == off) {
            pv->second += len;
            return;
        }
    }
    c.free[off] = len;
}

size_t Arena::trim() {
    size_t freed = 0;
    for (auto& c : chunks_) {
        if (!c.base || c.live) continue;
        (void)hipFree(c.base);
        c.base = nullptr;
        c.free.clear();
        reserved_ -= c.size;
        freed += c.size;
        c.size = 0;
    }
    return freed;
}

// Largest float x with sqrtf(x) < t (-1 when no x >= 0 qualifies).
// IEEE sqrtf is correctly rounded on the host and in the kernels (built
// with correctly rounded divide/sqrt) and monotone, so for any float d2,
// sqrtf(d2) < t holds exactly when d2 <= sqrt_below_bound(t).
float sqrt_below_bound(float t) {
    if (!(t > 0.0f) || !std::isfinite(t)) return t > 0.0f ? FLT_MAX : -1.0f;
    float x = t * t;
    while (x > 0.0f && !(std::sqrt(x) < t)) x = std::nextafter(x, 0.0f);
    for (;;) {
        const float y = std::nextafter(x, INFINITY);
        if (!(std::sqrt(y) < t)) break;
        x = y;
    }
    return x;
}

// ---------------------------------------------------------------------------
// Engine
// ---------------------------------------------------------------------------
Engine::Engine(int device, size_t budget) : device_(device), arena_(budget) {
    if (hipSetDevice(device) != hipSuccess) throw Error{SOP_E_HIP, "hipSetDevice failed"};
    if (hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking) != hipSuccess)
        throw Error{SOP_E_HIP, "hipStreamCreate failed"};
}

Engine::~Engine() {
    (void)hipSetDevice(device_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    drop_graph();
    for (auto& p : ev_pool_) {
        (void)hipEventDestroy(p.first);
        (void)hipEventDestroy(p.second);
    }
    for (auto& j : jobs_)
        for (void* b : j.blocks) arena_.release(b);
    (void)hipFree(d_jobs_);
    (void)hipFree(d_state_);
    (void)hipFree(d_tiles_);
    (void)hipFree(d_tiles1_);
    (void)hipFree(d_units_);
    (void)hipFree(d_units1_);
    (void)hipFree(d_sunits_);
    (void)hipFree(d_tea_tiles_);
    (void)hipFree(d_segs_);
    (void)hipFree(d_sunits1_);
    (void)hipFree(d_segs1_);
    (void)hipFree(d_scratch4_);
    if (aux_stream_) {
        (void)hipStreamSynchronize(aux_stream_);
        (void)hipStreamDestroy(aux_stream_);
    }
    if (ev_f_) (void)hipEventDestroy(ev_f_);
    for (hipEvent_t ev : ev_t2_)
        if (ev) (void)hipEventDestroy(ev);
    if (stream_) (void)hipStreamDestroy(stream_);
}

int Engine::check(hipError_t e, const char* what) {
    if (e == hipSuccess) return 0;
    return fail(SOP_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

HostJob* Engine::find(int id) {
    for (auto& j : jobs_)
        if (j.id == id) return &j;
    return nullptr;
}

int Engine::slot_of(int id) const {
    for (size_t i = 0; i < jobs_.size(); ++i)
        if (jobs_[i].id == id) return (int)i;
    return -1;
}

// Host-side construction of one job's maps, as the reference's
// constructors build them (covalent.cu:65-109, native.cu:58-107,
// pairlist.cu:41-57, pulling.cu:36-146, indentation.cu:84-200).
struct BuiltJob {
    int n = 0;
    int cov_cap = 1, nat_cap = 1, pair_cap = 1, pp_cap = 1, excl_cap = 1, pp_ld = 64;
    std::vector<int> cov_cnt, cov_j, nat_cnt, nat_j, pp_cnt, pp_j, excl_cnt, excl_j, pulled;
    std::vector<float> cov_r0;
    std::vector<float2> nat_par;
    std::vector<float4> coord, ext;
    std::vector<uint4> rng;
    DevJob dj{};
    DevJobState st{};
    long long end_step = 0;
};

static int build_job(const sop_params& p, const sop_system& s, long long numsteps, BuiltJob& b) {
    const int n = s.n;
    if (n <= 0 || !s.coords4) return fail(SOP_E_ARG, "system has no beads");
    if (p.timestep <= 0.0f || p.zeta <= 0.0f) return fail(SOP_E_ARG, "timestep and zeta must be positive");
    if (p.pairs_freq <= 0 || p.possiblepairs_freq <= 0 || p.nav <= 0 || p.outputtiming <= 0 || p.dcdfreq <= 0)
        return fail(SOP_E_ARG, "update frequencies must be positive");
    if (p.max_covalent <= 0 || p.max_native <= 0 || p.max_pairs <= 0 || p.max_possiblePairs <= 0)
        return fail(SOP_E_ARG, "list limits must be positive");
    b.n = n;
    auto in_range = [n](int i) { return i >= 0 && i < n; };
    // covalent
    b.cov_cnt.assign(n, 0);
    std::vector<std::vector<std::pair<int, float>>> cov(n);
    for (int k = 0; k < s.nbonds; ++k) {
        int i = s.bond_i[k], j = s.bond_j[k];
        if (!in_range(i) || !in_range(j)) return fail(SOP_E_ARG, "bond index out of range");
        cov[i].push_back({j, s.bond_r0[k]});
        cov[j].push_back({i, s.bond_r0[k]});
        if ((int)cov[i].size() > p.max_covalent || (int)cov[j].size() > p.max_covalent)
            return fail(SOP_E_LIMIT, "Maximum number of covalent bonds exceeded the limit of " +
                                         std::to_string(p.max_covalent) + ".");
    }
    for (int i = 0; i < n; ++i) b.cov_cap = std::max(b.cov_cap, (int)cov[i].size());
    b.cov_j.assign((size_t)b.cov_cap * n, 0);
    b.cov_r0.assign((size_t)b.cov_cap * n, 0.0f);
    for (int i = 0; i < n; ++i) {
        b.cov_cnt[i] = (int)cov[i].size();
        for (size_t k = 0; k < cov[i].size(); ++k) {
            b.cov_j[k * n + i] = cov[i][k].first;
            b.cov_r0[k * n + i] = cov[i][k].second;
        }
    }
    // native: r0^2 and -12 eh / r0^2 formed in double, stored as float
    b.nat_cnt.assign(n, 0);
    std::vector<std::vector<std::pair<int, float2>>> nat(n);
    for (int k = 0; k < s.nnative; ++k) {
        int i = s.nat_i[k], j = s.nat_j[k];
        if (!in_range(i) || !in_range(j)) return fail(SOP_E_ARG, "native contact index out of range");
        float r0 = s.nat_r0[k], eh = s.nat_eh[k];
        float2 par;
        par.x = r0 * r0;
        par.y = (float)(-12.0 * (double)eh / (double)(r0 * r0));
        nat[i].push_back({j, par});
        nat[j].push_back({i, par});
        if ((int)nat[i].size() > p.max_native || (int)nat[j].size() > p.max_native)
            return fail(SOP_E_LIMIT, "ERROR: Maximum number of native contacts exceeded the limit of " +
                                         std::to_string(p.max_native) + ".");
    }
    for (int i = 0; i < n; ++i) b.nat_cap = std::max(b.nat_cap, (int)nat[i].size());
    b.nat_j.assign((size_t)b.nat_cap * n, 0);
    b.nat_par.assign((size_t)b.nat_cap * n, float2{0.0f, 0.0f});
    for (int i = 0; i < n; ++i) {
        b.nat_cnt[i] = (int)nat[i].size();
        for (size_t k = 0; k < nat[i].size(); ++k) {
            b.nat_j[k * n + i] = nat[i][k].first;
            b.nat_par[k * n + i] = nat[i][k].second;
        }
    }
    // exclusions for the possible-pairs build: sorted unique partners
    b.excl_cnt.assign(n, 0);
    std::vector<std::vector<int>> ex(n);
    for (int i = 0; i < n; ++i) {
        for (auto& c : cov[i]) ex[i].push_back(c.first);
        for (auto& c : nat[i]) ex[i].push_back(c.first);
        std::sort(ex[i].begin(), ex[i].end());
        ex[i].erase(std::unique(ex[i].begin(), ex[i].end()), ex[i].end());
        b.excl_cap = std::max(b.excl_cap, (int)ex[i].size());
    }
    b.excl_j.assign((size_t)b.excl_cap * n, 0);
    for (int i = 0; i < n; ++i) {
        b.excl_cnt[i] = (int)ex[i].size();
        for (size_t k = 0; k < ex[i].size(); ++k) b.excl_j[k * n + i] = ex[i][k];
    }
    // possible list capacity: no bead has more than n-1 partners
    const int nm1 = std::max(1, n - 1);
    b.pp_cap = std::min(p.max_possiblePairs, nm1);
    b.pair_cap = std::min(p.max_pairs, nm1);
    b.pp_cnt.assign(n, 0);
    std::vector<std::vector<int>> pp(n);
    for (int k = 0; k < s.npairs; ++k) {
        int i = s.pair_i[k], j = s.pair_j[k];
        if (!in_range(i) || !in_range(j)) return fail(SOP_E_ARG, "pair index out of range");
        pp[i].push_back(j);
        pp[j].push_back(i);
        if ((int)pp[i].size() > p.max_possiblePairs || (int)pp[j].size() > p.max_possiblePairs)
            return fail(SOP_E_LIMIT, "ERROR: Maximum number of possible pairs exceeded the limit of " +
                                         std::to_string(p.max_possiblePairs) + ".");
    }
    for (int i = 0; i < n; ++i) b.pp_cap = std::max(b.pp_cap, (int)pp[i].size());
    b.pp_ld = (b.pp_cap + 63) / 64 * 64;  // rows start on 128-byte lines
    b.pp_j.assign((size_t)b.pp_ld * n, 0);
    for (int i = 0; i < n; ++i) {
        b.pp_cnt[i] = (int)pp[i].size();
        for (size_t k = 0; k < pp[i].size(); ++k) b.pp_j[(size_t)i * b.pp_ld + k] = pp[i][k];
    }
    // coordinates, w = 0 (gsop.cu:196)
    b.coord.resize(n);
    for (int i = 0; i < n; ++i)
        b.coord[i] = make_float4(s.coords4[4 * i], s.coords4[4 * i + 1], s.coords4[4 * i + 2], 0.0f);
    // RNG seeds (langevin.cu:23-24)
    b.rng.resize(n);
    if (s.rng_state4)
        std::memcpy(b.rng.data(), s.rng_state4, (size_t)n * 16);
    else
        generate_seeds(p.seed, n, reinterpret_cast<uint32_t*>(b.rng.data()));

    DevJob& J = b.dj;
    DevJobState& S = b.st;
    J.n = n;
    J.idx16 = n <= 65535 ? 1 : 0;
    J.cov_cap = b.cov_cap;
    J.nat_cap = b.nat_cap;
    J.pair_cap = b.pair_cap;
    J.pp_cap = b.pp_cap;
    J.pp_ld = b.pp_ld;
    J.excl_cap = b.excl_cap;
    J.ks = p.kspring_cov;
    J.R2 = p.R_limit * p.R_limit;
    J.a2 = p.a * p.a;
    J.m6 = -6.0f * p.el / J.a2;
    J.rc2 = p.pairs_cutoff * p.pairs_cutoff;
    J.el = p.el;
    J.pl_cut = p.pairlist_cutoff;
    J.pp_thr = p.pairs_threshold;
    J.pl_d2 = sqrt_below_bound(J.pl_cut);
    J.pp_d2 = sqrt_below_bound(J.pp_thr);
    J.h = p.timestep;
    J.zeta = p.zeta;
    J.hOverZeta = p.timestep / p.zeta;
    J.tempNorm = (float)((double)p.zeta / (6.0 * (double)p.timestep));
    J.nav = p.nav;
    J.pairs_freq = p.pairs_freq;
    J.pp_freq = p.possiblepairs_freq;
    if (p.minimization) J.flags |= JF_MINIMIZATION;
    if (p.heating) {
        if (p.tempFreq <= 0) return fail(SOP_E_ARG, "tempFreq must be positive");
        J.flags |= JF_HEATING;
        J.tempFreq = p.tempFreq;
        J.initialT = p.initialT;
        J.deltaT = p.deltaT;
    }
    {
        const float temp = p.heating ? p.initialT : p.temperature;
        S.var = sqrtf(2.0f * temp * p.zeta / p.timestep);
    }
    // pulling (pulling.cu:36-146)
    b.ext.assign(n, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    if (p.pulling) {
        if (p.deltax == 0.0f && p.fconst == 0.0f)
            return fail(SOP_E_CONFIG, "ERROR: Either 'deltax' or 'fconst' parameter should be specified to initiate pulling");
        J.flags |= JF_PULLING;
        if (!in_range(p.pulledEnd) || !in_range(p.fixedEnd)) return fail(SOP_E_ARG, "pulledEnd/fixedEnd out of range");
        float v[3];
        if (p.pullDirectionVector) {
            v[0] = p.pullVector[0];
            v[1] = p.pullVector[1];
            v[2] = p.pullVector[2];
        } else {
            const float4 a = b.coord[p.pulledEnd], c = b.coord[p.fixedEnd];
            float x = a.x - c.x, y = a.y - c.y, z = a.z - c.z;
            float l = sqrtf(x * x + y * y + z * z);
            v[0] = x / l;
            v[1] = y / l;
            v[2] = z / l;
        }
        for (int d = 0; d < 3; ++d) J.pullVec[d] = v[d];
        const float4 pe = b.coord[p.pulledEnd];
        J.cant0[0] = S.cant[0] = pe.x;
        J.cant0[1] = S.cant[1] = pe.y;
        J.cant0[2] = S.cant[2] = pe.z;
        for (int k = 0; k < s.npulled; ++k) {
            int i = s.pulled[k];
            // the reference accepts i == N (pulling.cu:113); it would write
            // past the system, so it is rejected here
            if (!in_range(i)) return fail(SOP_E_CONFIG, "Invalid pulling bead #" + std::to_string(i));
            b.ext[i] = make_float4(v[0] * p.fconst, v[1] * p.fconst, v[2] * p.fconst, 2.0f);
            b.pulled.push_back(i);
        }
        for (int k = 0; k < s.nfixed; ++k) {
            int i = s.fixed[k];
            if (!in_range(i)) return fail(SOP_E_CONFIG, "Invalid fixed bead #" + std::to_string(i));
            b.ext[i] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
        }
        if (p.deltax != 0.0f) {
            if (p.pullFreq <= 0) return fail(SOP_E_ARG, "pullFreq must be positive");
            J.flags |= JF_PULL_UPDATER;
            J.k_trans = p.k_trans;
            J.deltax = p.deltax;
            J.pullFreq = p.pullFreq;
            J.pulledEnd = p.pulledEnd;
        }
    }
    // indentation, continuous surface (indentation.cu:84-200)
    if (p.indentation) {
        J.flags |= JF_INDENTATION;
        for (int d = 0; d < 3; ++d) {
            S.chip[d] = J.chip0[d] = p.indChip[d];
            S.tip[d] = p.indTip[d];
            J.dir[d] = p.indDirection[d];
            S.micaR[d] = J.micaR0[d] = p.indSurfR0[d];
            J.micaN[d] = p.indSurfN[d];
        }
        S.V = p.indDeltaX;
        S.dx = 0.0f;
        J.tipR = p.indTipR;
        J.indKs = p.indKs;
        J.tipZeta = p.indTipZeta;
        J.moveSurface = p.indMoveSurface;
        J.fixTrans = p.indFixTrans;
        J.retraction = p.indRetractionStep;
        J.tipAp = 12.0f * p.indTipEl * powf(p.indTipSigma, 12.0f) * p.indTipA;
        J.tipBp = 6.0f * p.indTipEl * powf(p.indTipSigma, 6.0f) * p.indTipB;
        J.surfAp = 12.0f * p.indSurfEl * powf(p.indSurfSigma, 12.0f) * p.indSurfA;
        J.surfBp = 6.0f * p.indSurfEl * powf(p.indSurfSigma, 6.0f) * p.indSurfB;
        float norm = sqrtf(J.dir[0] * J.dir[0] + J.dir[1] * J.dir[1] + J.dir[2] * J.dir[2]);
        for (int d = 0; d < 3; ++d) J.dir[d] /= norm;
        norm = sqrtf(J.micaN[0] * J.micaN[0] + J.micaN[1] * J.micaN[1] + J.micaN[2] * J.micaN[2]);
        for (int d = 0; d < 3; ++d) J.micaN[d] /= norm;
    }
    // hydrodynamics (bdhitea.cu:30-100)
    if (p.hi_on) {
        if (p.minimization)
            return fail(SOP_E_CONFIG, "Using minimization with BDHI integrator is not supported and will never be.");
        J.flags |= JF_HI;
        J.hi_a = p.hi_a;
        if (p.hi_exact) {
            J.flags |= JF_HI_EXACT;
            if (p.hi_faithful_blocks) J.flags |= JF_HI_FAITHFUL;
            J.hi_chol_freq = p.hi_chol_freq > 0 ? p.hi_chol_freq : 1;
        } else {
            if (p.hi_epsilon_freq <= 0) return fail(SOP_E_CONFIG, "hi_epsilon_freq must be positive");
            J.hi_eps_freq = p.hi_epsilon_freq;
            J.hi_capricious = p.hi_capricious;
            J.hi_epsmax = p.hi_epsmax;
            if (!p.hi_unlisted) J.flags |= JF_HI_LISTED;
        }
    }
    // stride = min(nav, every created updater's frequency) (gsop.cu:116-122)
    int st = p.nav;
    auto mn = [&st](int f) { st = std::min(st, f); };
    if (p.indentation) mn(p.nav);
    if (J.flags & JF_PULL_UPDATER) mn(p.pullFreq);
    mn(p.possiblepairs_freq);
    mn(p.outputtiming);
    mn(p.pairs_freq);
    mn(p.dcdfreq);
    if (p.heating) mn(p.tempFreq);
    if (p.hi_on && !p.hi_exact) mn(p.hi_epsilon_freq);
    if (p.updater_freq_min < 0) return fail(SOP_E_ARG, "updater frequencies must be positive");
    if (p.updater_freq_min > 0) mn(p.updater_freq_min);  // updaters run by the host
    if (p.host_forces) J.flags |= JF_EXTRA_FORCE;
    J.stride = st;
    if (numsteps < 0) return fail(SOP_E_ARG, "numsteps must be >= 0");
    // the outer loop runs i = 0 .. numsteps/stride inclusive (gsop.cu:129)
    b.end_step = (numsteps / st + 1) * (long long)st;
    J.end_step = b.end_step;
    S.step = 0;
    return 0;
}

template <typename T>
static size_t nbytes(const std::vector<T>& v) {
    return v.size() * sizeof(T);
}

static unsigned long long round_up(unsigned long long x) { return (x + kAlign - 1) / kAlign * kAlign; }

static unsigned long long built_bytes(const BuiltJob& b, bool pulling, bool indentation) {
    const unsigned long long n = b.n;
    unsigned long long t = 0;
    t += 2 * round_up(n * 16);            // coord[2]
    t += round_up(n * 16);                // rng
    t += round_up(n * 4);                 // T
    if (pulling) t += round_up(n * 16) + round_up(n);  // ext, flags
    if (indentation) t += round_up(n * 16);
    t += round_up(n * 4) * 6;             // five count arrays, Verlet row order
    t += round_up(n * b.cov_cap * 4) * 2; // cov j, r0
    t += round_up(n * b.nat_cap * 4) + round_up(n * b.nat_cap * 8);
    if (b.dj.idx16) t += round_up(n * b.cov_cap * 2) + round_up(n * b.nat_cap * 2);  // 16-bit copies
    if (n <= (unsigned long long)kLdsStepBeads && !(b.dj.flags & JF_HI)) {  // bonded lists in row order
        const unsigned long long ld = (n + 63) / 64 * 64;
        t += round_up(n * 4) + round_up(ld * b.cov_cap * 6) + round_up(ld * b.nat_cap * 10);
    }
    const unsigned long long es = b.dj.idx16 ? 2 : 4;
    t += round_up((n + 63) / 64 * 64 * b.pair_cap * es) + round_up(n * b.pp_ld * es) + round_up(n * b.excl_cap * 4);
    t += round_up(std::max<size_t>(1, b.pulled.size()) * 4);
    t += round_up(n * 12);                // energies
    if (b.dj.flags & JF_EXTRA_FORCE) t += round_up(n * 16);  // forces of host potentials
    if (n >= (unsigned long long)kCellMinBeads && n <= 64ull * kMaxCellWords)
        t += round_up((n + 64) * 4) + round_up((n + 65) * 4) + 3 * round_up(n * 4);  // cell grid
    if (b.dj.flags & JF_HI) {
        t += 3 * round_up(n * 16) + round_up(n * 4);  // rforce, mforce, ci, eps
        if (b.dj.flags & JF_HI_EXACT) {
            const unsigned long long mp = (3 * n + 63) / 64 * 64;
            t += round_up(n * 16) + round_up(mp * mp * 4) + round_up(kCholLinvBytes) + round_up(chol_bbuf_bytes(mp)) +
                 2 * round_up(chol_panel_bytes(mp));
        }
    }
    return t;
}

int Engine::job_bytes(const sop_params& p, const sop_system& s, unsigned long long* bytes) {
    BuiltJob b;
    int rc = build_job(p, s, 0, b);
    if (rc) return rc;
    *bytes = built_bytes(b, p.pulling != 0, p.indentation != 0);
    return 0;
}

int Engine::add_job(const sop_params& p, const sop_system& s, long long numsteps, int* id) {
    if (hipSetDevice(device_) != hipSuccess) return fail(SOP_E_HIP, "hipSetDevice failed");
    BuiltJob b;
    int rc = build_job(p, s, numsteps, b);
    if (rc) return rc;
    HostJob hj;
    hj.n = b.n;
    hj.bytes = built_bytes(b, p.pulling != 0, p.indentation != 0);
    bool oom = false;
    auto A = [&](size_t bytes) -> void* {
        void* q = arena_.alloc(bytes);
        if (!q) {
            oom = true;
        } else {
            hj.blocks.push_back(q);
            hj.sizes.push_back(bytes);
        }
        return q;
    };
    DevJob& J = b.dj;
    const size_t n = b.n;
    J.coord[0] = (float4*)A(n * 16);
    J.coord[1] = (float4*)A(n * 16);
    J.rng = (uint4*)A(n * 16);
    J.T = (float*)A(n * 4);
    J.ext = p.pulling ? (float4*)A(n * 16) : nullptr;
    J.bflag = p.pulling ? (unsigned char*)A(n) : nullptr;
    J.tipF = p.indentation ? (float4*)A(n * 16) : nullptr;
    J.fextra = (J.flags & JF_EXTRA_FORCE) ? (float4*)A(n * 16) : nullptr;
    J.cov_cnt = (int*)A(n * 4);
    J.cov_j = (int*)A(n * b.cov_cap * 4);
    J.cov_j16 = J.idx16 ? (unsigned short*)A(n * b.cov_cap * 2) : nullptr;
    J.cov_r0 = (float*)A(n * b.cov_cap * 4);
    J.nat_cnt = (int*)A(n * 4);
    J.nat_j = (int*)A(n * b.nat_cap * 4);
    J.nat_j16 = J.idx16 ? (unsigned short*)A(n * b.nat_cap * 2) : nullptr;
    const bool rows_eligible = n <= (size_t)kLdsStepBeads && !(J.flags & JF_HI);  // may have its rows sorted
    const size_t rld = (n + 63) / 64 * 64;
    J.rb_cnt = rows_eligible ? (int*)A(n * 4) : nullptr;
    J.rb_cov_j = rows_eligible ? (unsigned short*)A(rld * b.cov_cap * 2) : nullptr;
    J.rb_cov_r0 = rows_eligible ? (float*)A(rld * b.cov_cap * 4) : nullptr;
    J.rb_nat_j = rows_eligible ? (unsigned short*)A(rld * b.nat_cap * 2) : nullptr;
    J.rb_nat_par = rows_eligible ? (float2*)A(rld * b.nat_cap * 8) : nullptr;
    J.nat_par = (float2*)A(n * b.nat_cap * 8);
    J.pair_cnt = (int*)A(n * 4);
    J.perm = (int*)A(n * 4);
    const size_t es = J.idx16 ? 2 : 4;
    J.pair_ld = (int)((n + 63) / 64 * 64);
    J.pair_j = A((size_t)J.pair_ld * b.pair_cap * es);
    J.pp_cnt = (int*)A(n * 4);
    J.pp_j = A(n * b.pp_ld * es);
    J.excl_cnt = (int*)A(n * 4);
    J.excl_j = (int*)A(n * b.excl_cap * 4);
    J.pulled = (int*)A(std::max<size_t>(1, b.pulled.size()) * 4);
    J.npulled = (int)b.pulled.size();
    J.energies = (float*)A(n * 12);
    J.lds_step = use_lds_step_ && n <= (size_t)kLdsStepBeads ? 1 : 0;
    J.sort_rows = J.lds_step && J.rb_cnt ? 1 : 0;
    if (uses_cells(n)) {
        J.cell_cap = (int)n + 64;
        J.cell_cnt = (int*)A((size_t)J.cell_cap * 4);
        J.cell_start = (int*)A((size_t)(J.cell_cap + 1) * 4);
        J.cell_beads = (int*)A(n * 4);
        J.bead_cell = (int*)A(n * 4);
        J.bead_slot = (int*)A(n * 4);
    }
    HiExactWork hw{};
    if (J.flags & JF_HI) {
        J.rforce = (float4*)A(n * 16);
        J.mforce = (float4*)A(n * 16);
        J.ci = (float4*)A(n * 16);
        J.eps = (float*)A(n * 4);
        if (J.flags & JF_HI_EXACT) {
            J.fdf = (float4*)A(n * 16);
            const size_t mp = (3 * n + 63) / 64 * 64;
            J.chol_ld = mp;
            J.chol_M = (float*)A(mp * mp * 4);
            hw.M = J.chol_M;
            hw.ld = mp;
            hw.mp = (int)mp;
            hw.mode = exact_mode_;
            hw.work.linv = (double*)A(kCholLinvBytes);
            hw.work.bbuf = (float*)A(chol_bbuf_bytes(mp));
            hw.work.phi = (unsigned short*)A(chol_panel_bytes(mp));
            hw.work.plo = (unsigned short*)A(chol_panel_bytes(mp));
            hw.chol_freq = J.hi_chol_freq;
            if (!aux_stream_) {  // the factorisation's lookahead stream, shared by the exact-HI jobs
                if (hipStreamCreateWithFlags(&aux_stream_, hipStreamNonBlocking) != hipSuccess ||
                    hipEventCreateWithFlags(&ev_f_, hipEventDisableTiming) != hipSuccess ||
                    hipEventCreateWithFlags(&ev_t2_[0], hipEventDisableTiming) != hipSuccess ||
                    hipEventCreateWithFlags(&ev_t2_[1], hipEventDisableTiming) != hipSuccess)
                    oom = true;
            }
            hw.aux = aux_stream_;
            hw.ev_f = ev_f_;
            hw.ev_t2[0] = ev_t2_[0];
            hw.ev_t2[1] = ev_t2_[1];
        }
    }
    if (oom) {
        for (void* q : hj.blocks) arena_.release(q);
        return fail(SOP_E_MEM, "HBM budget exhausted admitting a job of " + std::to_string(b.n) + " beads");
    }
    hipStream_t s0 = stream_;
    int err = 0;
    auto H2D = [&](void* dst, const void* src, size_t bytes) {
        if (!err && bytes) err = check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s0), "upload");
    };
    H2D(J.coord[0], b.coord.data(), nbytes(b.coord));
    H2D(J.coord[1], b.coord.data(), nbytes(b.coord));
    H2D(J.rng, b.rng.data(), nbytes(b.rng));
    std::vector<float> zeros(n * 4, 0.0f);
    H2D(J.T, zeros.data(), n * 4);
    if (J.ext) H2D(J.ext, b.ext.data(), nbytes(b.ext));
    std::vector<unsigned char> flags;
    if (J.bflag) {
        flags.resize(n);
        for (size_t i = 0; i < (size_t)n; ++i) flags[i] = (unsigned char)b.ext[i].w;
        H2D(J.bflag, flags.data(), flags.size());
    }
    if (J.tipF) H2D(J.tipF, zeros.data(), n * 16);
    if (J.fextra) H2D(J.fextra, zeros.data(), n * 16);
    H2D(J.cov_cnt, b.cov_cnt.data(), nbytes(b.cov_cnt));
    H2D(J.cov_j, b.cov_j.data(), nbytes(b.cov_j));
    std::vector<unsigned short> cov16, nat16;
    if (J.cov_j16) {
        cov16.assign(b.cov_j.begin(), b.cov_j.end());
        H2D(J.cov_j16, cov16.data(), nbytes(cov16));
    }
    H2D(J.cov_r0, b.cov_r0.data(), nbytes(b.cov_r0));
    H2D(J.nat_cnt, b.nat_cnt.data(), nbytes(b.nat_cnt));
    H2D(J.nat_j, b.nat_j.data(), nbytes(b.nat_j));
    if (J.nat_j16) {
        nat16.assign(b.nat_j.begin(), b.nat_j.end());
        H2D(J.nat_j16, nat16.data(), nbytes(nat16));
    }
    H2D(J.nat_par, b.nat_par.data(), nbytes(b.nat_par));
    // the row-ordered bonded lists for the identity row order
    std::vector<int> rbc;
    std::vector<unsigned short> rbcj, rbnj;
    std::vector<float> rbcr;
    std::vector<float2> rbnp;
    if (J.rb_cnt) {
        rbc.resize(n);
        rbcj.assign(rld * b.cov_cap, 0);
        rbcr.assign(rld * b.cov_cap, 0.0f);
        rbnj.assign(rld * b.nat_cap, 0);
        rbnp.assign(rld * b.nat_cap, float2{0.0f, 0.0f});
        for (size_t i = 0; i < n; ++i) {
            rbc[i] = b.cov_cnt[i] | (b.nat_cnt[i] << 16);
            for (int k = 0; k < b.cov_cap; ++k) {
                rbcj[k * rld + i] = (unsigned short)b.cov_j[k * n + i];
                rbcr[k * rld + i] = b.cov_r0[k * n + i];
            }
            for (int k = 0; k < b.nat_cap; ++k) {
                rbnj[k * rld + i] = (unsigned short)b.nat_j[k * n + i];
                rbnp[k * rld + i] = b.nat_par[k * n + i];
            }
        }
        H2D(J.rb_cnt, rbc.data(), nbytes(rbc));
        H2D(J.rb_cov_j, rbcj.data(), nbytes(rbcj));
        H2D(J.rb_cov_r0, rbcr.data(), nbytes(rbcr));
        H2D(J.rb_nat_j, rbnj.data(), nbytes(rbnj));
        H2D(J.rb_nat_par, rbnp.data(), nbytes(rbnp));
    }
    H2D(J.pair_cnt, zeros.data(), n * 4);
    std::vector<int> ident(n);
    for (size_t k = 0; k < n; ++k) ident[k] = (int)k;
    H2D(J.perm, ident.data(), n * 4);
    H2D(J.pp_cnt, b.pp_cnt.data(), nbytes(b.pp_cnt));
    std::vector<unsigned short> pp16;
    if (J.idx16) {
        pp16.assign(b.pp_j.begin(), b.pp_j.end());
        H2D(J.pp_j, pp16.data(), nbytes(pp16));
    } else {
        H2D(J.pp_j, b.pp_j.data(), nbytes(b.pp_j));
    }
    H2D(J.excl_cnt, b.excl_cnt.data(), nbytes(b.excl_cnt));
    H2D(J.excl_j, b.excl_j.data(), nbytes(b.excl_j));
    if (!b.pulled.empty()) H2D(J.pulled, b.pulled.data(), nbytes(b.pulled));
    if (J.chol_M) {
        // RPY matrix storage: the padding beyond 3n is an identity block the
        // factorisation leaves in place
        const size_t mp = J.chol_ld;
        if (!err) err = check(hipMemsetAsync(J.chol_M, 0, mp * mp * 4, s0), "matrix clear");
        std::vector<float> one(1, 1.0f);
        for (size_t r = 3 * n; r < mp && !err; ++r)
            err = check(hipMemcpyAsync(J.chol_M + r * mp + r, one.data(), 4, hipMemcpyHostToDevice, s0), "pad");
        if (!err) err = check(hipStreamSynchronize(s0), "pad sync");
    }
    if (!err) err = check(hipStreamSynchronize(s0), "upload sync");
    if (err) {
        for (void* q : hj.blocks) arena_.release(q);
        return err;
    }
    hj.id = next_id_++;
    hj.hw = hw;
    hj.dj = J;
    hj.st0 = b.st;
    hj.step = 0;
    hj.end_step = b.end_step;
    jobs_.push_back(hj);
    tables_dirty_ = true;
    drop_graph();
    // pre-loop Verlet build of runGPU (gsop.cu:126): from the [ pairs ]
    // section, at step 0
    rc = sync_tables();
    if (!rc) rc = rebuild_lists(hj.id, 2);
    if (rc) {
        remove_job(hj.id);
        return rc;
    }
    *id = hj.id;
    return 0;
}

int Engine::remove_job(int id) {
    int slot = slot_of(id);
    if (slot < 0) return fail(SOP_E_ARG, "no such job");
    if (int rc = pull_states()) return rc;
    for (void* b : jobs_[slot].blocks) arena_.release(b);
    jobs_.erase(jobs_.begin() + slot);
    tables_dirty_ = true;
    drop_graph();
    return 0;
}

// Copies the device state of the jobs that were on the device at the last
// table upload back into their host records, so a table rebuild (which
// moves jobs between slots) carries it over.
int Engine::pull_states() {
    if (synced_jobs_ <= 0 || !d_state_) return 0;
    if (int rc = synchronize()) return rc;
    std::vector<DevJobState> st(synced_jobs_);
    if (int rc = check(hipMemcpy(st.data(), d_state_, st.size() * sizeof(DevJobState), hipMemcpyDeviceToHost),
                       "state read"))
        return rc;
    for (int i = 0; i < synced_jobs_ && i < (int)jobs_.size(); ++i) jobs_[i].st0 = st[i];
    synced_jobs_ = 0;
    return 0;
}

// Uploads the job and tile tables when the job set changed.  Device state
// of surviving jobs is carried over through st0 (remove_job refreshed it).
int Engine::sync_tables() {
    if (!tables_dirty_) return 0;
    if (int rc = pull_states()) return rc;
    const int nj = (int)jobs_.size();
    if (nj > cap_jobs_) {
        int cap = std::max(16, nj * 2);
        (void)hipFree(d_jobs_);
        (void)hipFree(d_state_);
        d_jobs_ = nullptr;
        d_state_ = nullptr;
        if (int rc = check(hipMalloc(&d_jobs_, cap * sizeof(DevJob)), "job table")) return rc;
        if (int rc = check(hipMalloc(&d_state_, cap * sizeof(DevJobState)), "state table")) return rc;
        cap_jobs_ = cap;
    }
    std::vector<DevJob> dj(nj);
    std::vector<DevJobState> ds(nj);
    std::vector<int2> tiles;
    int g = 0;
    for (int k = 0; k < nj; ++k) {
        dj[k] = jobs_[k].dj;
        ds[k] = jobs_[k].st0;
        ds[k].step = jobs_[k].step;
        jobs_[k].tile0 = (int)tiles.size();
        for (int t = 0; t < jobs_[k].n; t += 64) tiles.push_back(make_int2(k, t));
        jobs_[k].ntile = (int)tiles.size() - jobs_[k].tile0;
        if (k == 0 || jobs_[k].dj.stride) g = std::gcd(g, jobs_[k].dj.stride);
    }
    L_ = g > 0 ? g : 1;
    any_tea_ = any_exact_ = false;
    cell_words_ = 0;
    any_idx32_ = 0;
    std::vector<int2> tea_tiles;  // the tiles of the TEA jobs
    for (auto& j : jobs_) {
        if (!j.dj.idx16) any_idx32_ = 1;
        cell_words_ = std::max(cell_words_, cell_words_of(j.dj));
        if ((j.dj.flags & JF_HI) && !(j.dj.flags & JF_HI_EXACT)) {
            any_tea_ = true;
            tea_tiles.insert(tea_tiles.end(), tiles.begin() + j.tile0, tiles.begin() + j.tile0 + j.ntile);
        }
        if (j.dj.flags & JF_HI_EXACT) any_exact_ = true;
    }
    ntea_tiles_ = (int)tea_tiles.size();
    if (int rc = check(upload_grow(d_tea_tiles_, cap_tea_tiles_, tea_tiles, stream_), "TEA tiles")) return rc;
    ntiles_ = (int)tiles.size();
    if (ntiles_ > cap_tiles_) {
        int cap = std::max(256, ntiles_ * 2);
        (void)hipFree(d_tiles_);
        d_tiles_ = nullptr;
        if (int rc = check(hipMalloc(&d_tiles_, cap * sizeof(int2)), "tile table")) return rc;
        cap_tiles_ = cap;
    }
    if (nj) {
        if (int rc = check(hipMemcpyAsync(d_jobs_, dj.data(), nj * sizeof(DevJob), hipMemcpyHostToDevice, stream_),
                           "job table upload"))
            return rc;
        if (int rc = check(hipMemcpyAsync(d_state_, ds.data(), nj * sizeof(DevJobState), hipMemcpyHostToDevice, stream_),
                           "state upload"))
            return rc;
    }
    if (ntiles_) {
        if (int rc = check(hipMemcpyAsync(d_tiles_, tiles.data(), ntiles_ * sizeof(int2), hipMemcpyHostToDevice, stream_),
                           "tile upload"))
            return rc;
    }
    // units of the LDS step kernel: (job slot, first bead) per kLdsUnit beads
    std::vector<int2> units;
    lds_beads_ = 0;
    any_global_ = 0;
    for (int k = 0; k < nj; ++k) {
        const DevJob& J = jobs_[k].dj;
        if (!J.lds_step) {
            any_global_ = 1;
            continue;
        }
        for (int b = 0; b < J.n; b += kLdsUnit) units.push_back(make_int2(k, b));
    }
    nunits_ = (int)units.size();
    if (nunits_ > cap_units_) {
        (void)hipFree(d_units_);
        d_units_ = nullptr;
        cap_units_ = 0;
        const int cap = std::max(64, nunits_ * 2);
        if (int rc = check(hipMalloc(&d_units_, cap * sizeof(int2)), "unit table")) return rc;
        cap_units_ = cap;
    }
    if (nunits_) {
        if (int rc = check(hipMemcpyAsync(d_units_, units.data(), nunits_ * sizeof(int2), hipMemcpyHostToDevice, stream_),
                           "unit upload"))
            return rc;
    }
    // LDS step plan
    std::vector<int2> su;
    std::vector<int4> sg;
    lds_beads_ = plan_lds_units(su, sg);
    nsunits_ = (int)su.size();
    if (int rc = check(upload_grow(d_sunits_, cap_sunits_, su, stream_), "step units")) return rc;
    if (int rc = check(upload_grow(d_segs_, cap_segs_, sg, stream_), "step segments")) return rc;
    if (int rc = check(hipStreamSynchronize(stream_), "table sync")) return rc;
    tables_dirty_ = false;
    synced_jobs_ = nj;
    return 0;
}

// The LDS step plan of the LDS-step jobs: each whole block of unit_rows
// rows of a job is a unit; the rows left over (or a whole small job) are
// packed, largest first, into units of at most unit_rows threads (segments
// rounded up to 64 rows) and kLdsStepBeads staged beads -- or, without
// shared units, each gets a unit of its own.  Returns the largest number of
// beads a unit stages.
int Engine::plan_lds_units(std::vector<int2>& units, std::vector<int4>& segs) const {
    struct Part {
        int job, row, rows, n;
    };
    std::vector<Part> parts;
    int lds = 0;
    units.clear();
    segs.clear();
    const int U = shape_.unit_rows;
    for (int k = 0; k < (int)jobs_.size(); ++k) {
        const DevJob& J = jobs_[k].dj;
        if (!J.lds_step) continue;
        int r = 0;
        for (; r + U <= J.n; r += U) {
            units.push_back(make_int2((int)segs.size(), 1));
            segs.push_back(make_int4(k, r, U, 0));
            lds = std::max(lds, J.n);
        }
        if (r < J.n) parts.push_back({k, r, J.n - r, J.n});
    }
    std::stable_sort(parts.begin(), parts.end(), [](const Part& a, const Part& b) { return a.rows > b.rows; });
    struct Bin {
        int threads, beads;
        std::vector<int> items;
    };
    std::vector<Bin> bins;
    for (int q = 0; q < (int)parts.size(); ++q) {
        const int span = (parts[q].rows + 63) / 64 * 64;
        Bin* fit = nullptr;
        for (auto& b : bins)
            if (shared_units_ && b.threads + span <= U && b.beads + parts[q].n <= kLdsStepBeads) {
                fit = &b;
                break;
            }
        if (!fit) {
            bins.push_back({0, 0, {}});
            fit = &bins.back();
        }
        fit->threads += span;
        fit->beads += parts[q].n;
        fit->items.push_back(q);
    }
    for (auto& b : bins) {
        units.push_back(make_int2((int)segs.size(), (int)b.items.size()));
        int off = 0;
        for (int q : b.items) {
            segs.push_back(make_int4(parts[q].job, parts[q].row, parts[q].rows, off));
            off += parts[q].n;
        }
        lds = std::max(lds, off);
    }
    return lds;
}

void Engine::drop_graph() {
    if (graphs_.empty()) return;
    // a replay of these graphs may still run on the stream: let it finish
    // before the executable graphs are destroyed
    (void)hipStreamSynchronize(stream_);
    for (auto& kv : graphs_)
        if (kv.second) (void)hipGraphExecDestroy(kv.second);
    graphs_.clear();
}

bool Engine::uses_cells(size_t n) const {
    return cell_min_ > 0 && n >= (size_t)cell_min_ && n <= (size_t)64 * kMaxCellWords;
}

int Engine::cell_words_of(const DevJob& J) const { return (use_cells_ && J.cell_cnt) ? (J.n + 63) / 64 : 0; }

void Engine::set_cells(int min_beads, int enabled) {
    if (min_beads != 0) cell_min_ = min_beads;
    use_cells_ = enabled != 0;
    tables_dirty_ = true;
    drop_graph();
}

void Engine::set_lds_step(int enabled) {
    use_lds_step_ = enabled != 0;
    for (auto& j : jobs_) {
        j.dj.lds_step = use_lds_step_ && j.n <= kLdsStepBeads ? 1 : 0;
        j.dj.sort_rows = j.dj.lds_step && j.dj.rb_cnt ? 1 : 0;
    }
    tables_dirty_ = true;
    drop_graph();
}

void Engine::set_tuning(int max_blocks, int use_graph) {
    max_blocks_ = max_blocks < 0 ? 0 : max_blocks;
    use_graph_ = use_graph != 0;
    drop_graph();
}

void Engine::profile(bool on) {
    if (!on) flush_profile();
    profiling_ = on;
    drop_graph();
}

void Engine::flush_profile() {
    if (ev_used_ == 0) return;
    (void)hipStreamSynchronize(stream_);
    for (size_t k = 0; k < ev_used_; ++k) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, ev_pool_[k].first, ev_pool_[k].second) == hipSuccess) {
            prof_ms_ += ms;
            prof_launches_ += ev_launches_[k];
        }
    }
    ev_used_ = 0;
}

void Engine::profile_read(long long* launches, double* ms) {
    flush_profile();
    *launches = prof_launches_;
    *ms = prof_ms_;
    prof_launches_ = 0;
    prof_ms_ = 0.0;
}

// The list kernels a block starting at the jobs' current steps has work for
// (the host mirrors every job's step): bit 0 the possible-pairs build, bit 1
// the Verlet build, bit 2 the Verlet row sort.  The kernels test the same
// cadence per job, so leaving out the others only drops launches that would
// return at once.
int Engine::lists_due(int slot_only) const {
    int m = 0;
    for (int q = 0; q < (int)jobs_.size(); ++q) {
        if (slot_only >= 0 && q != slot_only) continue;
        const HostJob& hj = jobs_[q];
        const long long s = hj.step;
        if (s >= hj.end_step || s % hj.dj.stride != 0) continue;
        if (s % hj.dj.pp_freq == 0) m |= 1;
        if (s % hj.dj.pairs_freq == 0) m |= 2;
        if (hj.dj.sort_rows && s % ((long long)hj.dj.pairs_freq * shape_.sort_every) == 0) m |= 4;
    }
    return m;
}

// Whether exact-HI job hj refactorises at step k of the block it is in.
static bool refactors(const HostJob& hj, int k) {
    const long long sk = hj.step + k;
    return sk < hj.end_step && sk % hj.hw.chol_freq == 0;
}

// What the next block of the batch launches beyond its fixed kernels: the
// list kernels due, then for each exact-HI job (slot order) the steps of the
// block that refactorise, each job's list closed by -1.  Blocks with equal
// keys launch the same kernels with the same arguments.
std::vector<long long> Engine::block_key() const {
    std::vector<long long> key{lists_due(-1)};
    if (any_exact_)
        for (const HostJob& hj : jobs_) {
            if (!(hj.dj.flags & JF_HI_EXACT)) continue;
            for (int k = 0; k < L_; ++k)
                if (refactors(hj, k)) key.push_back(k);
            key.push_back(-1);
        }
    return key;
}

// One block: per-job updaters, list rebuilds, L steps, step advance.
// slot_only >= 0 runs that job alone with its own stride (plain mode).
int Engine::launch_block(int slot_only) {
    const DevJob* jobs = d_jobs_;
    DevJobState* st = d_state_;
    const int2* tiles = d_tiles_;
    int nj = (int)jobs_.size(), nt = ntiles_, L = L_;
    if (slot_only >= 0) {
        if (int rc = tiles_for_slot(slot_only, &nt)) return rc;
        jobs = d_jobs_ + slot_only;
        st = d_state_ + slot_only;
        tiles = d_tiles1_;
        nj = 1;
        L = jobs_[slot_only].dj.stride;
    }
    if (int rc = check(launch_job_update(jobs, st, nj, stream_), "job update")) return rc;
    const int due = lists_due(slot_only);
    if (due & 1) {
        const hipError_t e =
            slot_only >= 0
                ? launch_possible(jobs, st, tiles, nt, nj, cell_words_of(jobs_[slot_only].dj),
                                  !jobs_[slot_only].dj.lds_step, d_sunits1_, d_segs1_,
                                  jobs_[slot_only].dj.lds_step ? nsunits1_ : 0, jobs_[slot_only].n, 0, stream_,
                                  shape_)
                : launch_possible(jobs, st, tiles, nt, nj, cell_words_, any_global_, d_sunits_, d_segs_, nsunits_,
                                  lds_beads_, 0, stream_, shape_);
        if (int rc = check(e, "possible pairs")) return rc;
    }
    if (due & 2) {
        const bool sort = (due & 4) != 0;
        const hipError_t e =
            slot_only >= 0
                ? launch_verlet(jobs, st, tiles, nt, !jobs_[slot_only].dj.idx16, !jobs_[slot_only].dj.lds_step,
                                d_units1_,
                                sort && jobs_[slot_only].dj.sort_rows ? (jobs_[slot_only].n + kLdsUnit - 1) / kLdsUnit
                                                                      : 0,
                                d_sunits1_, d_segs1_, jobs_[slot_only].dj.lds_step ? nsunits1_ : 0,
                                jobs_[slot_only].n, 0, stream_, shape_)
                : launch_verlet(jobs, st, tiles, nt, any_idx32_, any_global_, d_units_, sort ? nunits_ : 0,
                                d_sunits_, d_segs_, nsunits_, lds_beads_, 0, stream_, shape_);
        if (int rc = check(e, "verlet")) return rc;
    }
    const bool tea = slot_only >= 0 ? ((jobs_[slot_only].dj.flags & JF_HI) && !(jobs_[slot_only].dj.flags & JF_HI_EXACT))
                                    : any_tea_;
    const bool exact = slot_only >= 0 ? (jobs_[slot_only].dj.flags & JF_HI_EXACT) != 0 : any_exact_;
    if (tea)
        if (int rc = check(slot_only >= 0 ? launch_hi_tea_update(jobs, st, tiles, nt, nj, stream_)
                                          : launch_hi_tea_update(jobs, st, d_tea_tiles_, ntea_tiles_, nj, stream_),
                           "TEA update"))
            return rc;
    // profiling: HIP events around the step launches; when nothing else is
    // launched between them, one pair spans the block's L launches
    const bool span = !tea && !exact;
    for (int k = 0; k < L; ++k) {
        const bool prof = profiling_ && (!span || k == 0);
        const bool prof_end = profiling_ && (!span || k == L - 1);
        if (prof) {
            if (ev_used_ == ev_pool_.size()) {
                if (ev_pool_.size() >= 4096) flush_profile();
                if (ev_used_ == ev_pool_.size()) {
                    hipEvent_t a, b;
                    if (int rc = check(hipEventCreate(&a), "event")) return rc;
                    if (int rc = check(hipEventCreate(&b), "event")) return rc;
                    ev_pool_.push_back({a, b});
                    ev_launches_.push_back(1);
                }
            }
            (void)hipEventRecord(ev_pool_[ev_used_].first, stream_);
        }
        if (int rc = check(slot_only >= 0
                               ? launch_step(jobs, st, tiles, nt, !jobs_[slot_only].dj.lds_step, d_sunits1_, d_segs1_,
                                             jobs_[slot_only].dj.lds_step ? nsunits1_ : 0, jobs_[slot_only].n, k,
                                             max_blocks_, stream_, shape_)
                               : launch_step(jobs, st, tiles, nt, any_global_, d_sunits_, d_segs_, nsunits_, lds_beads_,
                                             k, max_blocks_, stream_, shape_),
                           "step"))
            return rc;
        if (prof_end) {
            (void)hipEventRecord(ev_pool_[ev_used_].second, stream_);
            ev_launches_[ev_used_] = span ? L : 1;
            ev_used_++;
        }
        if (tea)
            if (int rc = check(slot_only >= 0 ? launch_hi_tea_step(jobs, st, tiles, nt, k, stream_)
                                              : launch_hi_tea_step(jobs, st, d_tea_tiles_, ntea_tiles_, k, stream_),
                               "TEA step"))
                return rc;
        if (exact) {
            for (int q = 0; q < (int)jobs_.size(); ++q) {
                if (slot_only >= 0 && q != slot_only) continue;
                const HostJob& hj = jobs_[q];
                if (!(hj.dj.flags & JF_HI_EXACT)) continue;
                const int2* jt = slot_only >= 0 ? tiles : d_tiles_ + hj.tile0;
                const int jn = slot_only >= 0 ? nt : hj.ntile;
                const DevJob* tj = slot_only >= 0 ? jobs : d_jobs_;
                const DevJobState* ts = slot_only >= 0 ? st : d_state_;
                const bool refactor = refactors(hj, k);
                if (int rc = check(launch_chol_step(tj, ts, jt, jn, k, refactor ? &hj.hw : nullptr, stream_),
                                   "Cholesky step"))
                    return rc;
            }
        }
    }
    if (int rc = check(launch_job_advance(jobs, st, nj, L, stream_), "advance")) return rc;
    return 0;
}

// Tile list of one job (job index 0 relative to its slot), cached.
int Engine::tiles_for_slot(int slot, int* nt) {
    const int n = jobs_[slot].n;
    const int need = (n + 63) / 64;
    *nt = need;
    if (tiles1_n_ == n && tiles1_rows_ == shape_.unit_rows) return 0;
    if (need > cap_tiles1_) {
        (void)hipFree(d_tiles1_);
        d_tiles1_ = nullptr;
        cap_tiles1_ = 0;
        if (int rc = check(hipMalloc(&d_tiles1_, need * sizeof(int2)), "tile table")) return rc;
        cap_tiles1_ = need;
    }
    std::vector<int2> t1(need);
    for (int t = 0; t < need; ++t) t1[t] = make_int2(0, 64 * t);
    if (int rc = synchronize()) return rc;
    if (int rc = check(hipMemcpy(d_tiles1_, t1.data(), need * sizeof(int2), hipMemcpyHostToDevice), "tile upload"))
        return rc;
    const int nu = (n + kLdsUnit - 1) / kLdsUnit;
    if (nu > cap_units1_) {
        (void)hipFree(d_units1_);
        d_units1_ = nullptr;
        cap_units1_ = 0;
        if (int rc = check(hipMalloc(&d_units1_, nu * sizeof(int2)), "unit table")) return rc;
        cap_units1_ = nu;
    }
    std::vector<int2> u1(nu);
    for (int u = 0; u < nu; ++u) u1[u] = make_int2(0, kLdsUnit * u);  // row-sort units
    // step units of the job alone: unit_rows rows each, as in the batched plan
    const int U = shape_.unit_rows, ns = (n + U - 1) / U;
    std::vector<int2> su1(ns);
    std::vector<int4> sg1(ns);
    for (int u = 0; u < ns; ++u) {
        su1[u] = make_int2(u, 1);
        sg1[u] = make_int4(0, U * u, std::min(U, n - U * u), 0);
    }
    if (int rc = check(hipMemcpy(d_units1_, u1.data(), nu * sizeof(int2), hipMemcpyHostToDevice), "unit upload"))
        return rc;
    if (int rc = check(upload_grow(d_sunits1_, cap_sunits1_, su1, stream_), "step units")) return rc;
    if (int rc = check(upload_grow(d_segs1_, cap_segs1_, sg1, stream_), "step segments")) return rc;
    if (int rc = synchronize()) return rc;
    nsunits1_ = ns;
    tiles1_n_ = n;
    tiles1_rows_ = U;
    return 0;
}

int Engine::stage_job_update(int mask) {
    if (int rc = sync_tables()) return rc;
    return check(launch_job_update(d_jobs_, d_state_, (int)jobs_.size(), stream_, mask), "job update");
}
int Engine::stage_possible() {
    if (int rc = sync_tables()) return rc;
    return check(launch_possible(d_jobs_, d_state_, d_tiles_, ntiles_, (int)jobs_.size(), cell_words_, any_global_,
                                 d_sunits_,
                                 d_segs_, nsunits_, lds_beads_, 0, stream_, shape_),
                 "possible pairs");
}
int Engine::stage_verlet() {
    if (int rc = sync_tables()) return rc;
    return check(launch_verlet(d_jobs_, d_state_, d_tiles_, ntiles_, any_idx32_, any_global_, d_units_, nunits_,
                               d_sunits_, d_segs_, nsunits_, lds_beads_, 0, stream_, shape_),
                 "verlet");
}
int Engine::stage_tea_update() {
    if (int rc = sync_tables()) return rc;
    if (!any_tea_) return 0;
    return check(launch_hi_tea_update(d_jobs_, d_state_, d_tea_tiles_, ntea_tiles_, (int)jobs_.size(), stream_),
                 "TEA update");
}
int Engine::stage_step(int k) {
    if (int rc = sync_tables()) return rc;
    if (int rc = check(launch_step(d_jobs_, d_state_, d_tiles_, ntiles_, any_global_, d_sunits_, d_segs_, nsunits_,
                                   lds_beads_, k, max_blocks_, stream_, shape_),
                       "step"))
        return rc;
    if (any_tea_)
        if (int rc = check(launch_hi_tea_step(d_jobs_, d_state_, d_tea_tiles_, ntea_tiles_, k, stream_), "TEA step"))
            return rc;
    if (any_exact_)
        for (auto& hj : jobs_) {
            if (!(hj.dj.flags & JF_HI_EXACT)) continue;
            const bool refactor = refactors(hj, k);
            if (int rc = check(launch_chol_step(d_jobs_, d_state_, d_tiles_ + hj.tile0, hj.ntile, k,
                                                refactor ? &hj.hw : nullptr, stream_),
                               "Cholesky step"))
                return rc;
        }
    return 0;
}
int Engine::stage_advance(int nsteps) {
    if (int rc = sync_tables()) return rc;
    if (int rc = check(launch_job_advance(d_jobs_, d_state_, (int)jobs_.size(), nsteps, stream_), "advance")) return rc;
    for (auto& j : jobs_) j.step = std::min(j.step + nsteps, j.end_step);
    return 0;
}
int Engine::read_state_of(int id, DevJobState* out) {
    int slot = slot_of(id);
    if (slot < 0) return fail(SOP_E_ARG, "no such job");
    return read_state(slot, out);
}

int Engine::capture_graph(const std::vector<long long>& key) {
    hipGraph_t g = nullptr;
    if (int rc = check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "begin capture")) return rc;
    int rc = launch_block(-1);
    hipError_t e = hipStreamEndCapture(stream_, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (int r2 = check(e, "end capture")) return r2;
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return check(e, "graph instantiate");
    graphs_[key] = x;
    return 0;
}

int Engine::run_blocks(int nblocks) {
    if (hipSetDevice(device_) != hipSuccess) return fail(SOP_E_HIP, "hipSetDevice failed");
    if (int rc = sync_tables()) return rc;
    if (jobs_.empty()) return 0;
    // a block is replayed from the graph captured for its launch sequence;
    // past kMaxGraphs sequences (exact-HI jobs whose refactorisation steps
    // keep shifting against the block) the new ones are launched directly
    const bool graph = use_graph_ && !profiling_;
    for (int b = 0; b < nblocks; ++b) {
        const std::vector<long long> key = graph ? block_key() : std::vector<long long>();
        auto it = graph ? graphs_.find(key) : graphs_.end();
        if (graph && it == graphs_.end() && graphs_.size() < kMaxGraphs) {
            if (int rc = capture_graph(key)) return rc;
            it = graphs_.find(key);
        }
        if (it != graphs_.end()) {
            if (int rc = check(hipGraphLaunch(it->second, stream_), "graph launch")) return rc;
        } else {
            if (int rc = launch_block(-1)) return rc;
        }
        for (auto& j : jobs_) j.step = std::min(j.step + L_, j.end_step);
    }
    return 0;
}

int Engine::run_job(int id) {
    if (hipSetDevice(device_) != hipSuccess) return fail(SOP_E_HIP, "hipSetDevice failed");
    if (int rc = sync_tables()) return rc;
    int slot = slot_of(id);
    if (slot < 0) return fail(SOP_E_ARG, "no such job");
    HostJob& hj = jobs_[slot];
    while (hj.step < hj.end_step) {
        if (int rc = launch_block(slot)) return rc;
        hj.step = std::min(hj.step + hj.dj.stride, hj.end_step);
    }
    return 0;
}

int Engine::synchronize() {
    if (hipSetDevice(device_) != hipSuccess) return fail(SOP_E_HIP, "hipSetDevice failed");
    return check(hipStreamSynchronize(stream_), "synchronize");
}

int Engine::read_state(int slot, DevJobState* out) {
    if (int rc = sync_tables()) return rc;
    if (int rc = synchronize()) return rc;
    return check(hipMemcpy(out, d_state_ + slot, sizeof(DevJobState), hipMemcpyDeviceToHost), "state read");
}

int Engine::job_status(int id, long long* step, long long* end, int* overflow) {
    int slot = slot_of(id);
    if (slot < 0) return fail(SOP_E_ARG, "no such job");
    DevJobState s;
    if (int rc = read_state(slot, &s)) return rc;
    if (step) *step = s.step;
    if (end) *end = jobs_[slot].end_step;
    if (overflow) *overflow = s.overflow;
    return 0;
}

int Engine::get_indentation(int id, float* out) {
    int slot = slot_of(id);
    if (slot < 0) return fail(SOP_E_ARG, "no such job");
    DevJobState s;
    if (int rc = read_state(slot, &s)) return rc;
    for (int d = 0; d < 3; ++d) {
        out[d] = s.tip[d];
        out[3 + d] = s.chip[d];
        out[6 + d] = s.micaR[d];
    }
    for (int d = 0; d < 4; ++d) out[9 + d] = s.tipFav[d];
    return 0;
}

int Engine::get_coords(int id, float* out4) {
    int slot = slot_of(id);
    if (slot < 0) return fail(SOP_E_ARG, "no such job");
    DevJobState s;
    if (int rc = read_state(slot, &s)) return rc;
    const HostJob& hj = jobs_[slot];
    return check(hipMemcpy(out4, hj.dj.coord[s.step & 1], (size_t)hj.n * 16, hipMemcpyDeviceToHost), "coords read");
}

int Engine::get_coords_at(int id, long long s, float* out4) {
    HostJob* hj = find(id);
    if (!hj) return fail(SOP_E_ARG, "no such job");
    if (int rc = synchronize()) return rc;
    return check(hipMemcpy(out4, hj->dj.coord[s & 1], (size_t)hj->n * 16, hipMemcpyDeviceToHost), "coords read");
}

int Engine::upload_forces(int id, const float* f4) {
    HostJob* hj = find(id);
    if (!hj) return fail(SOP_E_ARG, "no such job");
    if (!hj->dj.fextra) return fail(SOP_E_STATE, "job takes no host forces (host_forces off)");
    // the last step that read the array cleared it; the copy is ordered
    // after that step and before the next on the engine's stream
    if (int rc = check(hipMemcpyAsync(hj->dj.fextra, f4, (size_t)hj->n * 16, hipMemcpyHostToDevice, stream_),
                       "forces upload"))
        return rc;
    return synchronize();
}

int Engine::set_coords(int id, const float* in4) {
    int slot = slot_of(id);
    if (slot < 0) return fail(SOP_E_ARG, "no such job");
    DevJobState s;
    if (int rc = read_state(slot, &s)) return rc;
    const HostJob& hj = jobs_[slot];
    return check(hipMemcpy(hj.dj.coord[s.step & 1], in4, (size_t)hj.n * 16, hipMemcpyHostToDevice), "coords write");
}

int Engine::get_array(int id, int what, void* out) {
    int slot = slot_of(id);
    if (slot < 0) return fail(SOP_E_ARG, "no such job");
    if (int rc = synchronize()) return rc;
    const HostJob& hj = jobs_[slot];
    const size_t n = hj.n;
    switch (what) {
    case 0: return check(hipMemcpy(out, hj.dj.T, n * 4, hipMemcpyDeviceToHost), "T read");
    case 1: return check(hipMemcpy(out, hj.dj.rng, n * 16, hipMemcpyDeviceToHost), "rng read");
    case 2:
        if (!hj.dj.ext) {
            std::memset(out, 0, n * 16);
            return 0;
        }
        return check(hipMemcpy(out, hj.dj.ext, n * 16, hipMemcpyDeviceToHost), "ext read");
    default: return fail(SOP_E_ARG, "unknown array");
    }
}

int Engine::list_cap(int id, int which, int* cap) {
    HostJob* hj = find(id);
    if (!hj) return fail(SOP_E_ARG, "no such job");
    switch (which) {
    case 0: *cap = hj->dj.cov_cap; break;
    case 1: *cap = hj->dj.nat_cap; break;
    case 2: *cap = hj->dj.pair_cap; break;
    case 3: *cap = hj->dj.pp_cap; break;
    default: return fail(SOP_E_ARG, "unknown list");
    }
    return 0;
}

int Engine::get_list(int id, int which, int* counts, int* entries) {
    HostJob* hj = find(id);
    if (!hj) return fail(SOP_E_ARG, "no such job");
    if (int rc = synchronize()) return rc;
    const DevJob& J = hj->dj;
    const size_t n = hj->n;
    const int* c;
    const void* e;
    int cap;
    bool typed = false;  // list words uint16/int32 per J.idx16
    switch (which) {
    case 0: c = J.cov_cnt; e = J.cov_j; cap = J.cov_cap; break;
    case 1: c = J.nat_cnt; e = J.nat_j; cap = J.nat_cap; break;
    case 2: c = J.pair_cnt; e = J.pair_j; cap = J.pair_cap; typed = true; break;
    case 3: c = J.pp_cnt; e = J.pp_j; cap = J.pp_cap; typed = true; break;
    default: return fail(SOP_E_ARG, "unknown list");
    }
    if (int rc = check(hipMemcpy(counts, c, n * 4, hipMemcpyDeviceToHost), "list counts")) return rc;
    // device layout: slot-major [k*ld + r] (Verlet: ld = pair_ld; covalent,
    // native: ld = n) or bead-major [i*ld + k] (possible: ld = pp_ld)
    const size_t ld = which == 2 ? (size_t)J.pair_ld : which == 3 ? (size_t)J.pp_ld : n;
    const size_t nw = which == 3 ? n * ld : ld * cap;
    std::vector<int> words(nw);
    if (typed && J.idx16) {
        std::vector<unsigned short> w16(nw);
        if (int rc = check(hipMemcpy(w16.data(), e, nw * 2, hipMemcpyDeviceToHost), "list entries")) return rc;
        words.assign(w16.begin(), w16.end());
    } else {
        if (int rc = check(hipMemcpy(words.data(), e, nw * 4, hipMemcpyDeviceToHost), "list entries")) return rc;
    }
    if (which == 2) {  // row r holds bead perm[r]'s list
        std::vector<int> perm(n), rc(counts, counts + n);
        if (int rv = check(hipMemcpy(perm.data(), J.perm, n * 4, hipMemcpyDeviceToHost), "row order")) return rv;
        for (size_t r = 0; r < n; ++r) {
            const size_t i = (size_t)perm[r];
            counts[i] = rc[r];
            for (int k = 0; k < cap; ++k) entries[(size_t)k * n + i] = words[(size_t)k * ld + r];
        }
        return 0;
    }
    if (which != 3) {
        std::memcpy(entries, words.data(), n * cap * 4);
        return 0;
    }
    // the possible list is kept bead-major on the device
    for (size_t i = 0; i < n; ++i)
        for (int k = 0; k < cap; ++k) entries[(size_t)k * n + i] = words[i * ld + k];
    return 0;
}

int Engine::get_hi(int id, double* eps, float* beta) {
    int slot = slot_of(id);
    if (slot < 0) return fail(SOP_E_ARG, "no such job");
    DevJobState s;
    if (int rc = read_state(slot, &s)) return rc;
    if (eps) *eps = s.eps_traj;
    if (beta) *beta = s.beta;
    return 0;
}

// The factor's leading 3n x 3n block (lower triangle; upper set to zero).
int Engine::get_hi_factor(int id, float* L) {
    HostJob* hj = find(id);
    if (!hj) return fail(SOP_E_ARG, "no such job");
    if (!hj->dj.chol_M) return fail(SOP_E_STATE, "job has no Cholesky factor (hi_exact off)");
    if (int rc = synchronize()) return rc;
    const size_t m = 3 * (size_t)hj->n, ld = hj->dj.chol_ld;
    std::vector<float> M(ld * ld);
    if (int rc = check(hipMemcpy(M.data(), hj->dj.chol_M, ld * ld * 4, hipMemcpyDeviceToHost), "factor read"))
        return rc;
    for (size_t r = 0; r < m; ++r)
        for (size_t c = 0; c < m; ++c) L[r * m + c] = c <= r ? M[r * ld + c] : 0.0f;
    return 0;
}

int Engine::compute_forces(int id, float* out4) {
    int slot = slot_of(id);
    if (slot < 0) return fail(SOP_E_ARG, "no such job");
    if (int rc = sync_tables()) return rc;
    const HostJob& hj = jobs_[slot];
    if ((size_t)hj.n > scratch4_n_) {
        (void)hipFree(d_scratch4_);
        d_scratch4_ = nullptr;
        if (int rc = check(hipMalloc(&d_scratch4_, (size_t)hj.n * 16), "scratch")) return rc;
        scratch4_n_ = hj.n;
    }
    // the step kernel the job is stepped by (LDS or global gathers), in its
    // forces-only instantiation, over the job's own tiles / units
    int nt = 0;
    if (int rc = tiles_for_slot(slot, &nt)) return rc;
    const bool lds = hj.dj.lds_step != 0;
    if (int rc = check(launch_forces(d_jobs_ + slot, d_state_ + slot, d_tiles1_, nt, !lds, d_sunits1_, d_segs1_,
                                     lds ? nsunits1_ : 0, hj.n, d_scratch4_, stream_, shape_),
                       "forces"))
        return rc;
    if (int rc = synchronize()) return rc;
    return check(hipMemcpy(out4, d_scratch4_, (size_t)hj.n * 16, hipMemcpyDeviceToHost), "forces read");
}

// SOPPotential::computeEnergy + sumEnergies (gsop.cu:255-271): per-bead
// energies, summed in double and halved because every pair is stored on
// both beads.
int Engine::compute_energies(int id, float* per_bead, float* sums3) {
    int slot = slot_of(id);
    if (slot < 0) return fail(SOP_E_ARG, "no such job");
    if (int rc = sync_tables()) return rc;
    const HostJob& hj = jobs_[slot];
    const size_t n = hj.n;
    if (int rc = check(launch_energies(d_jobs_, d_state_, slot, hj.n, stream_), "energies")) return rc;
    if (int rc = synchronize()) return rc;
    std::vector<float> e(3 * n);
    if (int rc = check(hipMemcpy(e.data(), hj.dj.energies, 3 * n * 4, hipMemcpyDeviceToHost), "energies read"))
        return rc;
    if (per_bead) std::memcpy(per_bead, e.data(), 3 * n * 4);
    if (sums3) {
        for (int p = 0; p < 3; ++p) {
            double acc = 0.0;
            for (size_t i = 0; i < n; ++i) acc += e[p * n + i];
            sums3[p] = (float)(acc * 0.5f);
        }
    }
    return 0;
}

int Engine::rebuild_lists(int id, int mask) {
    int slot = slot_of(id);
    if (slot < 0) return fail(SOP_E_ARG, "no such job");
    if (int rc = sync_tables()) return rc;
    int nt = 0;
    if (int rc = tiles_for_slot(slot, &nt)) return rc;
    if (mask & 1)
        if (int rc = check(launch_possible(d_jobs_ + slot, d_state_ + slot, d_tiles1_, nt, 1,
                                           cell_words_of(jobs_[slot].dj), !jobs_[slot].dj.lds_step, d_sunits1_,
                                           d_segs1_,
                                           jobs_[slot].dj.lds_step ? nsunits1_ : 0, jobs_[slot].n, 1, stream_,
                                           shape_),
                           "possible pairs"))
            return rc;
    if (mask & 2)
        if (int rc = check(launch_verlet(d_jobs_ + slot, d_state_ + slot, d_tiles1_, nt, !jobs_[slot].dj.idx16,
                                         !jobs_[slot].dj.lds_step,
                                         d_units1_,
                                         jobs_[slot].dj.sort_rows ? (jobs_[slot].n + kLdsUnit - 1) / kLdsUnit : 0,
                                         d_sunits1_, d_segs1_, jobs_[slot].dj.lds_step ? nsunits1_ : 0,
                                         jobs_[slot].n, 1, stream_, shape_),
                           "verlet"))
            return rc;
    return synchronize();
}

int Engine::list_stats(double* avg4, long long* beads) {
    if (int rc = synchronize()) return rc;
    double sum[4] = {0, 0, 0, 0};
    long long nb = 0;
    for (auto& hj : jobs_) {
        const int* cnts[4] = {hj.dj.cov_cnt, hj.dj.nat_cnt, hj.dj.pair_cnt, hj.dj.pp_cnt};
        std::vector<int> c(hj.n);
        for (int w = 0; w < 4; ++w) {
            if (int rc = check(hipMemcpy(c.data(), cnts[w], hj.n * 4, hipMemcpyDeviceToHost), "counts")) return rc;
            for (int v : c) sum[w] += v;
        }
        nb += hj.n;
    }
    for (int w = 0; w < 4; ++w) avg4[w] = nb ? sum[w] / nb : 0.0;
    *beads = nb;
    return 0;
}

Tuning Engine::tuning() const {
    return Tuning{max_blocks_,          shape_.unit_rows, shared_units_ ? 1 : 0, shape_.pair_batch,
                  shape_.verlet_chunk, shape_.sort_every, shape_.split_rows, shape_.split_global};
}

int Engine::set_shape(const Tuning& t) {
    const bool ok = t.max_blocks >= 0 && t.unit_rows >= 64 && t.unit_rows <= 1024 && t.unit_rows % 64 == 0 &&
                    (t.pair_batch == 4 || t.pair_batch == 6 || t.pair_batch == 8) &&
                    (t.verlet_chunk == 4 || t.verlet_chunk == 8) && t.sort_every >= 1 && t.sort_every <= 1000 &&
                    (t.split_rows == 0 || t.split_rows == 4 || t.split_rows == 8 || t.split_rows == 16) &&
                    (t.split_global == 0 || t.split_global == 4 || t.split_global == 8 || t.split_global == 16);
    if (!ok) return fail(SOP_E_ARG, "launch shape out of range");
    max_blocks_ = t.max_blocks;
    shape_.unit_rows = t.unit_rows;
    shape_.pair_batch = t.pair_batch;
    shape_.verlet_chunk = t.verlet_chunk;
    shape_.sort_every = t.sort_every;
    shape_.split_rows = t.split_rows;
    shape_.split_global = t.split_global;
    shared_units_ = t.shared_units != 0;
    tables_dirty_ = true;  // the unit plan depends on unit_rows and shared_units
    tiles1_n_ = -1;
    drop_graph();
    return 0;
}

// Every job's arena blocks and the state table, copied into one device
// buffer (allocated here; the caller frees it with hipFree).
int Engine::snapshot(void** out) {
    *out = nullptr;
    if (int rc = sync_tables()) return rc;
    if (int rc = synchronize()) return rc;
    size_t total = jobs_.size() * sizeof(DevJobState);
    for (auto& j : jobs_)
        for (size_t b : j.sizes) total += b;
    char* buf = nullptr;
    if (int rc = check(hipMalloc(&buf, total), "tuning snapshot")) return rc;
    size_t off = 0;
    int err = 0;
    for (auto& j : jobs_)
        for (size_t k = 0; k < j.blocks.size() && !err; ++k) {
            err = check(hipMemcpyAsync(buf + off, j.blocks[k], j.sizes[k], hipMemcpyDeviceToDevice, stream_), "snapshot");
            off += j.sizes[k];
        }
    if (!err)
        err = check(hipMemcpyAsync(buf + off, d_state_, jobs_.size() * sizeof(DevJobState), hipMemcpyDeviceToDevice,
                                   stream_),
                    "snapshot");
    if (!err) err = synchronize();
    if (err) {
        (void)hipFree(buf);
        return err;
    }
    *out = buf;
    return 0;
}

// Puts the snapshot back (arrays, state table, host step mirrors).
int Engine::restore(void* snap, const std::vector<long long>& steps) {
    if (int rc = synchronize()) return rc;
    if (int rc = sync_tables()) return rc;  // slots as at the snapshot: the job set is unchanged
    const char* buf = static_cast<const char*>(snap);
    size_t off = 0;
    for (auto& j : jobs_)
        for (size_t k = 0; k < j.blocks.size(); ++k) {
            if (int rc = check(hipMemcpyAsync(j.blocks[k], buf + off, j.sizes[k], hipMemcpyDeviceToDevice, stream_),
                               "restore"))
                return rc;
            off += j.sizes[k];
        }
    if (int rc = check(hipMemcpyAsync(d_state_, buf + off, jobs_.size() * sizeof(DevJobState), hipMemcpyDeviceToDevice,
                                      stream_),
                       "restore"))
        return rc;
    for (size_t q = 0; q < jobs_.size(); ++q) jobs_[q].step = steps[q];
    return synchronize();
}

// GPU time of nblocks eager blocks, from events on the engine's stream.
int Engine::time_blocks(int nblocks, double* ms) {
    if (int rc = sync_tables()) return rc;
    hipEvent_t a = nullptr, b = nullptr;
    if (int rc = check(hipEventCreate(&a), "event")) return rc;
    if (int rc = check(hipEventCreate(&b), "event")) {
        (void)hipEventDestroy(a);
        return rc;
    }
    int rc = check(hipEventRecord(a, stream_), "event");
    for (int k = 0; k < nblocks && !rc; ++k) {
        rc = launch_block(-1);
        for (auto& j : jobs_) j.step = std::min(j.step + L_, j.end_step);
    }
    if (!rc) rc = check(hipEventRecord(b, stream_), "event");
    if (!rc) rc = check(hipEventSynchronize(b), "event");
    float f = 0.0f;
    if (!rc) rc = check(hipEventElapsedTime(&f, a, b), "event");
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms = f;
    return rc;
}

// Coordinate descent over the launch shapes (each dimension in turn, the
// others held at their best so far), every candidate timed on the same
// window of real blocks from the same snapshot: whole periods of the jobs'
// list cadences (possible pairs, Verlet rebuilds, the longest row-sort
// cadence tried), so each candidate pays its share of every kernel.
int Engine::autotune(Tuning* best_out, double* ms_default, double* ms_best, int* trials_out) {
    if (hipSetDevice(device_) != hipSuccess) return fail(SOP_E_HIP, "hipSetDevice failed");
    if (int rc = sync_tables()) return rc;
    const Tuning start = tuning();
    if (best_out) *best_out = start;
    if (trials_out) *trials_out = 0;
    if (jobs_.empty()) return 0;
    bool lds = false;
    long long window = L_;
    for (auto& j : jobs_) {
        if (j.step >= j.end_step) continue;
        if (j.dj.lds_step) lds = true;
        window = std::max<long long>(window, j.dj.pp_freq);
        if (j.dj.sort_rows) window = std::max<long long>(window, (long long)j.dj.pairs_freq * 40);
    }
    window = std::min<long long>((window + L_ - 1) / L_ * L_, 4000);
    const int nblocks = (int)std::max<long long>(1, window / L_);
    std::vector<long long> steps;
    for (auto& j : jobs_) steps.push_back(j.step);
    const bool saved_graph = use_graph_, saved_prof = profiling_;
    profiling_ = false;
    void* snap = nullptr;
    if (int rc = snapshot(&snap)) return rc;
    int trials = 0;
    auto trial = [&](const Tuning& t, double* per_step) -> int {
        if (int rc = set_shape(t)) return rc;
        if (int rc = restore(snap, steps)) return rc;
        double ms = 0.0;
        if (int rc = time_blocks(nblocks, &ms)) return rc;
        *per_step = ms / ((double)nblocks * L_);
        ++trials;
        return 0;
    };
    Tuning best = start;
    double best_ms = 0.0;
    int rc = trial(best, &best_ms);
    const double def_ms = best_ms;
    auto consider = [&](Tuning t) {
        if (rc) return;
        double ms = 0.0;
        rc = trial(t, &ms);
        if (!rc && ms < best_ms * 0.995) {  // a real gain, beyond run-to-run noise
            best = t;
            best_ms = ms;
        }
    };
    if (lds) {
        // smaller units spread a few small jobs over more CUs
        for (int v : {1024, 768, 512, 256, 128})
            if (v != best.unit_rows) { Tuning t = best; t.unit_rows = v; consider(t); }
        { Tuning t = best; t.shared_units = !best.shared_units; consider(t); }
        for (int v : {4, 8})
            if (v != best.pair_batch) { Tuning t = best; t.pair_batch = v; consider(t); }
        { Tuning t = best; t.verlet_chunk = best.verlet_chunk == 8 ? 4 : 8; consider(t); }
        for (int v : {10, 40})
            if (v != best.sort_every) { Tuning t = best; t.sort_every = v; consider(t); }
        // a few small jobs: each bead's terms spread over a workgroup
        if (ntiles_ <= kSplitMaxTiles)
            for (int v : {16, 8, 4})
                if (v != best.split_rows) { Tuning t = best; t.split_rows = v; consider(t); }
    }
    if (any_global_) {
        const int full = (ntiles_ + 3) / 4;
        for (int b : {256, 1024, 4096})
            if (b < full && b != best.max_blocks) { Tuning t = best; t.max_blocks = b; consider(t); }
        if (best.max_blocks != 0) { Tuning t = best; t.max_blocks = 0; consider(t); }
        // each bead's terms spread over a workgroup: more wavefronts to hide the gathers
        for (int v : {16, 8, 4})
            if (v != best.split_global) { Tuning t = best; t.split_global = v; consider(t); }
    }
    // the jobs as they were, under the chosen shape
    const int rc2 = set_shape(rc ? start : best);
    const int rc3 = restore(snap, steps);
    (void)hipFree(snap);
    use_graph_ = saved_graph;
    profiling_ = saved_prof;
    if (rc) return rc;
    if (rc2) return rc2;
    if (rc3) return rc3;
    if (best_out) *best_out = best;
    if (ms_default) *ms_default = def_ms;
    if (ms_best) *ms_best = best_ms;
    if (trials_out) *trials_out = trials;
    return 0;
}

}  // namespace sop
