// synth.cpp -- seeded synthetic SOP systems for benchmarks and tests.
//
// A chain is a self-avoiding walk with 3.8 A bonds confined to a sphere
// whose volume gives `density` cubic angstroms per bead, i.e. a compact
// globule.  Native contacts are all pairs with |i-j| > 2 closer than 8 A,
// with well depths uniform in [eh_lo, eh_hi]; bond and contact lengths are
// the generated distances, so the structure is the native state.  The
// shell variant packs chains into the Voronoi cells of points spread
// over a sphere (a capsid-like layer) and adds inter-chain contacts.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "../../include/sopgpu.h"
#include "engine.h"

namespace {

struct Rng {  // xoshiro256** seeded through splitmix64
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        for (int i = 0; i < 4; ++i) {
            seed += 0x9e3779b97f4a7c15ull;
            uint64_t z = seed;
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
            s[i] = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uni() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct V3 {
    double x, y, z;
};

struct Grid {
    double cell;
    std::unordered_map<int64_t, std::vector<int>> m;
    explicit Grid(double c) : cell(c) {}
    int64_t key(int a, int b, int c) const {
        return ((int64_t)(a + 1048576) << 42) | ((int64_t)(b + 1048576) << 21) | (int64_t)(c + 1048576);
    }
    void cellof(const V3& p, int& a, int& b, int& c) const {
        a = (int)std::floor(p.x / cell);
        b = (int)std::floor(p.y / cell);
        c = (int)std::floor(p.z / cell);
    }
    void add(const V3& p, int id) {
        int a, b, c;
        cellof(p, a, b, c);
        m[key(a, b, c)].push_back(id);
    }
    void remove(const V3& p, int id) {
        int a, b, c;
        cellof(p, a, b, c);
        auto& v = m[key(a, b, c)];
        for (size_t k = 0; k < v.size(); ++k)
            if (v[k] == id) {
                v[k] = v.back();
                v.pop_back();
                return;
            }
    }
    template <typename F>
    void near(const V3& p, F f) const {
        int a, b, c;
        cellof(p, a, b, c);
        for (int da = -1; da <= 1; ++da)
            for (int db = -1; db <= 1; ++db)
                for (int dc = -1; dc <= 1; ++dc) {
                    auto it = m.find(key(a + da, b + db, c + dc));
                    if (it == m.end()) continue;
                    for (int id : it->second) f(id);
                }
    }
};

const double kBond = 3.8, kMin = 4.0, kContact = 8.0;

static double dist(const V3& a, const V3& b) {
    double x = a.x - b.x, y = a.y - b.y, z = a.z - b.z;
    return std::sqrt(x * x + y * y + z * z);
}

// Grows one confined self-avoiding walk of n beads starting at `start`,
// appending to pos; `inside` tells whether a point is allowed.  A bead that
// cannot be placed makes the walk back up a few beads and try again.
template <typename Inside>
static bool grow(int n, const V3& start, Rng& rng, Inside inside, std::vector<V3>& pos, Grid& g) {
    const int base = (int)pos.size();
    pos.push_back(start);
    g.add(start, base);
    double slack = 1.0;
    int backtracks = 0;
    while ((int)pos.size() - base < n) {
        const V3 prev = pos.back();
        bool ok = false;
        V3 q{};
        for (int attempt = 0; attempt < 600 && !ok; ++attempt) {
            double u = 2.0 * rng.uni() - 1.0, ph = 2.0 * M_PI * rng.uni(), s = std::sqrt(1.0 - u * u);
            q = V3{prev.x + kBond * s * std::cos(ph), prev.y + kBond * s * std::sin(ph), prev.z + kBond * u};
            if (!inside(q, slack)) continue;
            bool clash = false;
            const int self = (int)pos.size() - 1;
            g.near(q, [&](int id) {
                if (id != self && dist(pos[id], q) < kMin) clash = true;
            });
            ok = !clash;
        }
        if (ok) {
            pos.push_back(q);
            g.add(q, (int)pos.size() - 1);
            continue;
        }
        if (++backtracks > 20000) return false;
        if (backtracks % 200 == 0) slack *= 1.02;
        const int drop = std::min<int>(1 + backtracks % 8, (int)pos.size() - base - 1);
        for (int d = 0; d < drop; ++d) {
            g.remove(pos.back(), (int)pos.size() - 1);
            pos.pop_back();
        }
    }
    return true;
}

static void contacts(const std::vector<V3>& pos, const std::vector<int>& chain, Rng& rng, float lo, float hi,
                     std::vector<int>& ci, std::vector<int>& cj, std::vector<float>& r0, std::vector<float>& eh,
                     const std::vector<float>& fx) {
    Grid g(kContact);
    const int n = (int)pos.size();
    for (int i = 0; i < n; ++i) g.add(pos[i], i);
    std::vector<int> cand;
    for (int i = 0; i < n; ++i) {
        cand.clear();
        g.near(pos[i], [&](int j) {
            if (j > i) cand.push_back(j);
        });
        std::sort(cand.begin(), cand.end());
        for (int j : cand) {
            if (chain[i] == chain[j] && j - i <= 2) continue;
            double x = (double)fx[4 * j] - fx[4 * i], y = (double)fx[4 * j + 1] - fx[4 * i + 1],
                   z = (double)fx[4 * j + 2] - fx[4 * i + 2];
            double d = std::sqrt(x * x + y * y + z * z);
            if (d < kContact) {
                ci.push_back(i);
                cj.push_back(j);
                r0.push_back((float)d);
                eh.push_back((float)(lo + (hi - lo) * rng.uni()));
            }
        }
    }
}

template <typename T>
static T* dup(const std::vector<T>& v) {
    T* p = (T*)std::malloc(std::max<size_t>(1, v.size()) * sizeof(T));
    if (!v.empty()) std::memcpy(p, v.data(), v.size() * sizeof(T));
    return p;
}

static void emit(const std::vector<V3>& pos, const std::vector<int>& chain, Rng& rng, float lo, float hi,
                 sop_synth* out) {
    const int n = (int)pos.size();
    std::vector<float> c4(4 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        c4[4 * i] = (float)pos[i].x;
        c4[4 * i + 1] = (float)pos[i].y;
        c4[4 * i + 2] = (float)pos[i].z;
        c4[4 * i + 3] = 0.0f;
    }
    std::vector<int> bi, bj;
    std::vector<float> br;
    for (int i = 0; i + 1 < n; ++i)
        if (chain[i] == chain[i + 1]) {
            double x = (double)c4[4 * i + 4] - c4[4 * i], y = (double)c4[4 * i + 5] - c4[4 * i + 1],
                   z = (double)c4[4 * i + 6] - c4[4 * i + 2];
            bi.push_back(i);
            bj.push_back(i + 1);
            br.push_back((float)std::sqrt(x * x + y * y + z * z));
        }
    std::vector<int> ci, cj;
    std::vector<float> r0, eh;
    contacts(pos, chain, rng, lo, hi, ci, cj, r0, eh, c4);
    out->n = n;
    out->coords4 = dup(c4);
    out->nbonds = (int)bi.size();
    out->bond_i = dup(bi);
    out->bond_j = dup(bj);
    out->bond_r0 = dup(br);
    out->nnative = (int)ci.size();
    out->nat_i = dup(ci);
    out->nat_j = dup(cj);
    out->nat_r0 = dup(r0);
    out->nat_eh = dup(eh);
}

}  // namespace

extern "C" int sop_synth_chain(int n, unsigned long long seed, float density, float eh_lo, float eh_hi,
                               sop_synth* out) {
    if (n < 1 || !out || density <= 0.0f) return sop::fail(SOP_E_ARG, "sop_synth_chain: bad arguments");
    std::memset(out, 0, sizeof(*out));
    Rng rng(seed);
    double R = std::cbrt(3.0 * n * density / (4.0 * M_PI));
    for (int tries = 0; tries < 20; ++tries) {
        std::vector<V3> pos;
        Grid g(kMin);
        auto inside = [R](const V3& q, double slack) {
            return q.x * q.x + q.y * q.y + q.z * q.z <= R * R * slack * slack;
        };
        if (grow(n, V3{0, 0, 0}, rng, inside, pos, g)) {
            std::vector<int> chain(n, 0);
            emit(pos, chain, rng, eh_lo, eh_hi, out);
            return 0;
        }
        R *= 1.05;
    }
    return sop::fail(SOP_E_ARG, "sop_synth_chain: walk could not be grown");
}

extern "C" int sop_synth_shell(int nchains, int n, unsigned long long seed, float radius, float eh_lo, float eh_hi,
                               sop_synth* out) {
    if (nchains < 1 || n < 1 || radius <= 0.0f || !out) return sop::fail(SOP_E_ARG, "sop_synth_shell: bad arguments");
    std::memset(out, 0, sizeof(*out));
    Rng rng(seed);
    const double density = 175.0;
    const double thick = nchains * (double)n * density / (4.0 * M_PI * radius * radius);
    // chain centres: a Fibonacci lattice on the sphere with a seeded rotation
    std::vector<V3> ctr(nchains);
    const double ga = M_PI * (3.0 - std::sqrt(5.0)), rot = 2.0 * M_PI * rng.uni();
    for (int c = 0; c < nchains; ++c) {
        double z = 1.0 - 2.0 * (c + 0.5) / nchains, r = std::sqrt(1.0 - z * z), ph = ga * c + rot;
        ctr[c] = V3{radius * r * std::cos(ph), radius * r * std::sin(ph), radius * z};
    }
    std::vector<V3> pos;
    std::vector<int> chain;
    Grid g(kMin);
    for (int c = 0; c < nchains; ++c) {
        auto inside = [&, c](const V3& q, double slack) {
            double rq = std::sqrt(q.x * q.x + q.y * q.y + q.z * q.z);
            if (std::fabs(rq - radius) > 0.5 * thick * slack) return false;
            double dc = dist(q, ctr[c]);
            for (int o = 0; o < nchains; ++o)
                if (o != c && dist(q, ctr[o]) < dc / slack) return false;
            return true;
        };
        const size_t before = pos.size();
        if (!grow(n, ctr[c], rng, inside, pos, g)) return sop::fail(SOP_E_ARG, "sop_synth_shell: walk could not be grown");
        chain.resize(pos.size(), c);
        (void)before;
    }
    emit(pos, chain, rng, eh_lo, eh_hi, out);
    return 0;
}

extern "C" void sop_free_system(sop_synth* s) {
    if (!s) return;
    std::free(s->coords4);
    std::free(s->bond_i);
    std::free(s->bond_j);
    std::free(s->bond_r0);
    std::free(s->nat_i);
    std::free(s->nat_j);
    std::free(s->nat_r0);
    std::free(s->nat_eh);
    std::free(s->pair_i);
    std::free(s->pair_j);
    std::memset(s, 0, sizeof(*s));
}
