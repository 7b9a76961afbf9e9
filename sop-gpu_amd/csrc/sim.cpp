// sim.cpp -- the sop-gpu program on the MI355X engine.
//
// Parameter file -> .top -> coordinates -> plugins in initFF's order ->
// runGPU's loop -> outputs, as src/sop-gpu.cpp:18-88 and src/gsop.cu:45-163
// do.  The plugin interfaces keep the reference's shape
// (SOPPotential/SOPUpdater/SOPIntegrator, src/gsop.h:69-105): the
// registries' order fixes the order of the updater passes and of the
// energy columns.  Built-in potentials do their work inside the fused step
// kernel, so their compute() launches nothing; the integrator's integrate()
// launches the step.  Trajectories of the many-runs mode are separate jobs
// of one engine, seeded from one seed table as the reference seeds its
// concatenated bead array.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <memory>
#include <string>
#include <vector>

#include "../../include/sop_plugin.h"
#include "../../include/sopgpu.h"
#include "config.h"
#include "engine.h"
#include "hostutil.h"
#include "io.h"

namespace sop {

struct Sim {
    Config cfg;
    Topology top;
    std::unique_ptr<Engine> eng;
    std::vector<int> jobs;  // one per trajectory, admitted when the run starts
    int Ntr = 1, firstrun = 0, nav = 1000, N = 0, device = 0;
    std::string device_name;
    long long numsteps = 0, step = 0;
    int substep = 0;
    bool hi = false, tea = false, running = false, ran = false;
    time_t t_start = 0;
    sop_params p{};
    sop_sim* handle = nullptr;
    std::vector<std::unique_ptr<SOPPotential>> potentials;
    std::vector<std::unique_ptr<SOPUpdater>> updaters;
    std::unique_ptr<SOPIntegrator> integrator;
    size_t builtin_updaters = 0;  // the registry's first entries are the program's own
    // inputs of the trajectories (admitted by run())
    std::vector<std::vector<float>> xyz;
    std::vector<uint32_t> seeds;
    std::vector<int> fixed, pulled;
    // forces host potentials add in the current step, per trajectory
    std::vector<std::vector<float>> host_forces;
    std::vector<char> host_forces_set;
    // per-output-step caches
    long long energy_step = -1;
    std::vector<float> energies;  // [traj*3 + term]
    std::vector<Atom> extra_atoms;  // indentation representation beads
    std::vector<float> surface0;    // their surface grid at dx = 0
    float cant_vec[3] = {0, 0, 0};  // cantilever base offset from the chip

    void check(int rc) const {
        if (rc) throw ConfigError(sop_last_error());
    }
    // A file-name parameter: its masked value when the file sets it, else
    // the default as written -- defaults are not masked, so an unset
    // `outputname` names a file with a literal "<name>" in it, as in the
    // reference (getParameterAs with a default, src/IO/configreader.h:71-77).
    std::string file_param(const std::string& key, const std::string& def) const {
        return cfg.has(key) ? cfg.str(key) : def;
    }
    std::string run_name(const std::string& pattern, int traj) const {
        return string_replace(pattern, "<run>", std::to_string(traj + firstrun));
    }
    std::vector<float> coords(int traj) {
        std::vector<float> c(4 * (size_t)N);
        check(eng->get_coords(jobs[traj], c.data()));
        return c;
    }
    void compute_energies() {
        if (energy_step == step) return;
        energies.assign(3 * Ntr, 0.0f);
        for (int t = 0; t < Ntr; ++t) check(eng->compute_energies(jobs[t], nullptr, &energies[3 * t]));
        energy_step = step;
    }
};

// A plugin's failure message, set by sop_sim_fail() in a callback.
static thread_local std::string g_plugin_error;

// ---- host plugins (sop_sim_add_updater / sop_sim_add_potential) ----------
struct HostUpdater : SOPUpdater {
    sop_sim* h;
    sop_plugin_fn fn;
    void* user;
    sop_release_fn rel;
    HostUpdater(sop_sim* h_, const char* n, int f, sop_plugin_fn fn_, void* u, sop_release_fn r)
        : h(h_), fn(fn_), user(u), rel(r) {
        name = n;
        frequency = f;
    }
    ~HostUpdater() override {
        if (rel) rel(user);
    }
    void update() override {
        g_plugin_error.clear();
        if (fn(h, user))
            throw ConfigError(g_plugin_error.empty() ? "updater '" + name + "' failed" : g_plugin_error);
    }
};

struct HostPotential : SOPPotential {
    sop_sim* h;
    sop_plugin_fn fn;
    sop_energy_fn efn;
    void* user;
    sop_release_fn rel;
    int count;
    HostPotential(sop_sim* h_, const char* n, int c, sop_plugin_fn fn_, sop_energy_fn e, void* u, sop_release_fn r)
        : h(h_), fn(fn_), efn(e), user(u), rel(r), count(c) {
        name = n;
    }
    ~HostPotential() override {
        if (rel) rel(user);
    }
    void compute() override {
        g_plugin_error.clear();
        if (fn && fn(h, user))
            throw ConfigError(g_plugin_error.empty() ? "potential '" + name + "' failed" : g_plugin_error);
    }
    int getEnergiesCount() const override { return count; }
    float getEnergy(int traj, int id) override { return efn ? efn(h, traj, id, user) : 0.0f; }
};

// ---- potentials (creation order of initFF, gsop.cu:80-85) ----------------
// The built-in potentials compute inside the fused step kernel, so their
// compute() launches nothing.
struct TermPotential : SOPPotential {
    Sim& sim;
    int term;
    TermPotential(Sim& s, const char* n, int t) : sim(s), term(t) { name = n; }
    void compute() override {}
    int getEnergiesCount() const override { return 1; }
    float getEnergy(int traj, int) override { return sim.energies[3 * traj + term]; }
};

// Indentation and pulling add their forces in the step kernel and report
// no energy term (indentation.h, pulling.h).
struct KernelPotential : SOPPotential {
    explicit KernelPotential(const char* n) { name = n; }
    void compute() override {}
};

// ---- updaters --------------------------------------------------------------
struct IndentationTipUpdater : SOPUpdater {  // indentation.cu:414-550
    Sim& sim;
    int outputFreq;
    FILE* out = nullptr;
    float chipAv[3] = {0, 0, 0}, tipAv[3] = {0, 0, 0}, fav[4] = {0, 0, 0, 0}, kDeltaXAv = 0.0f;
    IndentationTipUpdater(Sim& s) : sim(s) {
        name = "Indentation";
        frequency = s.nav;
        outputFreq = (int)s.cfg.integer_or("indentationOutputFreq", 1000);
        const std::string fn = s.file_param("indentationOutput", "indentation.<name>_<run>.dat");
        out = std::fopen(fn.c_str(), "w");
        if (!out) throw ConfigError("Can't open '" + fn + "'");
    }
    ~IndentationTipUpdater() override {
        if (out) std::fclose(out);
    }
    void update() override {
        if (sim.step == sim.p.indRetractionStep) std::printf("Starting tip retraction at step %lld.\n", sim.step);
        sim.check(sim.eng->stage_job_update(1));
        if (sim.step % frequency != 0) return;
        DevJobState S;
        sim.check(sim.eng->read_state_of(sim.jobs[0], &S));
        const float Ks = sim.p.indKs;
        for (int d = 0; d < 3; ++d) {
            chipAv[d] += S.chip[d];
            tipAv[d] += S.tip[d];
            fav[d] += S.tipFav[d];
        }
        fav[3] += S.tipFav[3];
        kDeltaXAv += S.tipDr * Ks;
        if (sim.step % outputFreq == 0) {
            const int factor = outputFreq / frequency;
            for (int d = 0; d < 3; ++d) {
                chipAv[d] /= factor;
                tipAv[d] /= factor;
                fav[d] /= factor;
            }
            fav[3] /= factor;
            kDeltaXAv /= factor;
            const float fmod = sqrtf(fav[0] * fav[0] + fav[1] * fav[1] + fav[2] * fav[2]);
            const float dxav = (float)(S.dx - 0.5 * factor * S.V);
            std::printf("Indentation: %lld\t%5.2f\t%5.2f\t%5.2f\n", sim.step, dxav, fmod, kDeltaXAv);
            std::fprintf(out,
                         "%lld\t%8.5f\t%8.5f\t%8.5f\t%8.5f\t%8.5f\t%8.5f\t%8.5f\t%8.5f\t%8.5f\t%8.5f\t%8.5f\t%8.5f\t"
                         "%8.5f\t%8.5f\n",
                         sim.step, S.dx, fav[3], fmod, std::fabs(S.tipDr * Ks), kDeltaXAv, fav[0], fav[1], fav[2],
                         S.tip[0], S.tip[1], S.tip[2], S.chip[0], S.chip[1], S.chip[2]);
            std::fflush(out);
            for (int d = 0; d < 3; ++d) chipAv[d] = tipAv[d] = fav[d] = 0.0f;
            fav[3] = kDeltaXAv = 0.0f;
        }
    }
};

// Beads that show the cantilever base, the tip and the surface in the DCD
// and PDB output (indentation.cu:254-400), and the VMD script that connects
// the surface beads.
static void indentation_representation(Sim& sim) {
    const Config& c = sim.cfg;
    const int size = (int)c.integer_or("indentationSurfaceSize", 51);
    const float step = (float)c.real_or("indentationSurfaceStep", 1.4);
    const float cant = (float)c.real_or("indentationCantLength", 500.0);
    const sop_params& p = sim.p;
    std::vector<Atom>& ex = sim.extra_atoms;
    ex.assign(2 + (size_t)size * size, Atom());
    for (int k = 0; k < 2; ++k) {
        ex[k].id = sim.N + k;
        ex[k].resName = "TIP";
        ex[k].name = "CA";
        ex[k].chain = 'T';
        ex[k].resid = k + 1;
    }
    ex[0].x = p.indChip[0];
    ex[0].y = p.indChip[1];
    ex[0].z = p.indChip[2];
    ex[1].x = p.indTip[0];
    ex[1].y = p.indTip[1];
    ex[1].z = p.indTip[2];
    // a vector perpendicular to the indentation direction, of the cantilever's length
    const float* d = p.indDirection;
    float r[3];
    if (d[0] != 0.0f) {
        r[1] = 1.0f;
        r[2] = 0.0f;
        r[0] = -r[1] * d[1] / d[0];
        const float len = sqrtf(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        for (float& v : r) v /= len;
    } else {
        r[0] = 1.0f;
        r[1] = r[2] = 0.0f;
    }
    for (float& v : r) v *= cant;
    std::printf("Cantilever representation vector: (%5.3f, %5.3f, %5.3f)\n", r[0], r[1], r[2]);
    for (int k = 0; k < 3; ++k) sim.cant_vec[k] = r[k];
    ex[0].x += r[0];
    ex[0].y += r[1];
    ex[0].z += r[2];
    // the surface grid, in the plane through indentationSurfaceR0 normal to indentationSurfaceN
    const float* n = p.indSurfN;
    const float* r0 = p.indSurfR0;
    float r1[3];
    if (n[0] == 0) {
        r1[0] = 1.0f, r1[1] = 0, r1[2] = 0;
    } else if (n[1] == 0) {
        r1[0] = 0, r1[1] = 1.0f, r1[2] = 0;
    } else if (n[2] == 0) {
        r1[0] = 0, r1[1] = 0, r1[2] = 1.0f;
    } else {
        r1[0] = 0, r1[1] = 1.0f, r1[2] = -(n[1] / n[2]) * r1[1];
    }
    const float norm = (float)std::sqrt(r1[0] * r1[0] + r1[1] * r1[1] + r1[2] * r1[2]);
    for (float& v : r1) v /= norm;
    const float r2[3] = {n[1] * r1[2] - n[2] * r1[1], n[2] * r1[0] - n[0] * r1[2], n[0] * r1[1] - n[1] * r1[0]};
    sim.surface0.assign(3 * (size_t)size * size, 0.0f);
    for (int i = 0; i < size; ++i)
        for (int j = 0; j < size; ++j) {
            const int i1 = i - size / 2, i2 = j - size / 2;
            Atom& a = ex[2 + i * size + j];
            float q[3];
            for (int k = 0; k < 3; ++k) q[k] = r0[k] + i1 * step * r1[k] + i2 * step * r2[k];
            a.id = sim.N + 2 + i * size + j;
            a.resName = "MIC";
            a.name = "CA";
            a.chain = 'M';
            a.resid = 1;
            a.x = q[0];
            a.y = q[1];
            a.z = q[2];
            for (int k = 0; k < 3; ++k) sim.surface0[3 * (i * size + j) + k] = q[k];
        }
    const std::string vmd = c.str_or("indentationSurfConnectFile", "connect_mica.vmd");
    FILE* f = std::fopen(vmd.c_str(), "w");
    if (!f) throw ConfigError("Can't open '" + vmd + "'");
    std::printf("Dumping VMD script to connect surface elements into '%s'.\n", vmd.c_str());
    const int N = sim.N;
    std::fprintf(f, "set sel [atomselect top \"index %d\"]\n", N);
    std::fprintf(f, "set bonds {%d}\n", N + 1);
    std::fprintf(f, "$sel setbonds [list $bonds]\n$sel delete\n");
    std::fprintf(f, "set sel [atomselect top \"index %d\"]\n", N + 1);
    std::fprintf(f, "set bonds {%d}\n", N);
    std::fprintf(f, "$sel setbonds [list $bonds]\n$sel delete\n");
    for (int i = 0; i < size; ++i)
        for (int j = 0; j < size; ++j) {
            std::fprintf(f, "set sel [atomselect top \"index %d\"]\n", 2 + i * size + j + N);
            std::fprintf(f, "set bonds {");
            if (i != size - 1) std::fprintf(f, "%d", 2 + (i + 1) * size + j + N);
            if (j != size - 1) std::fprintf(f, " %d", 2 + i * size + j + 1 + N);
            std::fprintf(f, "}\n");
            std::fprintf(f, "$sel setbonds [list $bonds]\n$sel delete\n");
        }
    std::fclose(f);
}

struct IndentationAminoUpdater : SOPUpdater {  // indentation.cu:238-246, 555-580
    Sim& sim;
    IndentationAminoUpdater(Sim& s) : sim(s) {
        name = "AdditionalAminos";
        frequency = s.p.outputtiming;
    }
    void update() override {
        if (sim.step % frequency != 0) return;
        DevJobState S;
        sim.check(sim.eng->read_state_of(sim.jobs[0], &S));
        std::vector<Atom>& ex = sim.extra_atoms;
        ex[1].x = S.tip[0];
        ex[1].y = S.tip[1];
        ex[1].z = S.tip[2];
        if (sim.p.indMoveSurface) {
            const float* d = sim.p.indDirection;
            for (size_t i = 0; 2 + i < ex.size(); ++i) {
                ex[2 + i].x = sim.surface0[3 * i] + (float)S.dx * d[0];
                ex[2 + i].y = sim.surface0[3 * i + 1] + (float)S.dx * d[1];
                ex[2 + i].z = sim.surface0[3 * i + 2] + (float)S.dx * d[2];
            }
        } else {
            ex[0].x = S.chip[0] + sim.cant_vec[0];
            ex[0].y = S.chip[1] + sim.cant_vec[1];
            ex[0].z = S.chip[2] + sim.cant_vec[2];
        }
    }
};

struct PullingUpdater : SOPUpdater {  // pulling.cu:152-214
    Sim& sim;
    std::vector<std::string> files;
    PullingUpdater(Sim& s) : sim(s) {
        name = "Pulling";
        frequency = s.p.pullFreq;
        const std::string pattern = s.file_param("pullOutput", "pull.<name>_<run>.dat");
        for (int t = 0; t < s.Ntr; ++t) {
            files.push_back(s.run_name(pattern, t));
            FILE* f = std::fopen(files.back().c_str(), "w");
            if (!f) throw ConfigError("Can't open '" + files.back() + "'");
            std::fclose(f);
        }
    }
    void update() override {
        sim.check(sim.eng->stage_job_update(2));
        if (sim.step % frequency != 0) return;
        for (int t = 0; t < sim.Ntr; ++t) {
            DevJobState S;
            sim.check(sim.eng->read_state_of(sim.jobs[t], &S));
            const HostJob* hj = sim.eng->find(sim.jobs[t]);
            const std::vector<float> c = sim.coords(t);
            const int pe = sim.p.pulledEnd, fe = sim.p.fixedEnd;
            const float vx = c[4 * pe] - c[4 * fe], vy = c[4 * pe + 1] - c[4 * fe + 1], vz = c[4 * pe + 2] - c[4 * fe + 2];
            const float* pv = hj->dj.pullVec;
            const float ex = vx * pv[0] + vy * pv[1] + vz * pv[2];
            const float e = sqrtf(vx * vx + vy * vy + vz * vz);
            const float f = S.extF[0] * pv[0] + S.extF[1] * pv[1] + S.extF[2] * pv[2];
            if (t == 0 && sim.step % 100000 == 0) {
                std::printf("'Cantilever chip' coordinates for run #%d: %f, %f, %f\n", sim.firstrun, S.cant[0],
                            S.cant[1], S.cant[2]);
                std::printf("'Cantilever tip' coordinates for run #%d: %f, %f, %f\n", sim.firstrun, c[4 * pe],
                            c[4 * pe + 1], c[4 * pe + 2]);
            }
            FILE* out = std::fopen(files[t].c_str(), "a");
            if (!out) throw ConfigError("Can't append to '" + files[t] + "'");
            std::fprintf(out, "%12lld\t%5.3f\t%5.3f\t%5.3f\t%5.3f\t%5.3f\t%5.3f\n", sim.step, e, ex, f, S.extF[0],
                         S.extF[1], S.extF[2]);
            std::fclose(out);
        }
    }
};

struct PossiblepairUpdater : SOPUpdater {
    Sim& sim;
    PossiblepairUpdater(Sim& s) : sim(s) {
        name = "Possiblepairlist";
        frequency = s.p.possiblepairs_freq;
    }
    void update() override { sim.check(sim.eng->stage_possible()); }
};

struct PairlistUpdater : SOPUpdater {
    Sim& sim;
    PairlistUpdater(Sim& s) : sim(s) {
        name = "Pairlist";
        frequency = s.p.pairs_freq;
    }
    void update() override { sim.check(sim.eng->stage_verlet()); }
};

struct OutputManager : SOPUpdater {  // output_manager.cu:73-360
    Sim& sim;
    std::vector<std::string> files;
    int width, printRuns;
    bool computeRg;
    float R_limit, R_limit_bond;
    std::vector<float> temperatures, rgs;
    std::vector<int> nativeCounts;
    OutputManager(Sim& s) : sim(s) {
        name = "DAT output";
        frequency = s.p.outputtiming;
        width = (int)s.cfg.integer_or("outputcolwidth", 16);
        printRuns = (int)s.cfg.integer_or("printruns", 10);
        computeRg = s.cfg.boolean_or("computeRg", false);
        if (s.cfg.has("mode"))
            throw ConfigError("'mode' parameter is deprecated. Use 'computeRg' parameter to turn the Rg computation on/off.");
        R_limit = s.p.R_limit;
        R_limit_bond = s.p.R_limit_bond;
        const std::string pattern = s.file_param("outputname", "energy.<name>_<run>.dat");
        for (int t = 0; t < s.Ntr; ++t) {
            files.push_back(s.run_name(pattern, t));
            FILE* f = std::fopen(files.back().c_str(), "w");
            if (!f) throw ConfigError("Can't open '" + files.back() + "'");
            std::fclose(f);
        }
        temperatures.assign(s.Ntr, 0.0f);
        rgs.assign(s.Ntr, 0.0f);
        nativeCounts.assign(s.Ntr, 0);
    }
    void update() override {
        if (sim.step % frequency != 0) return;  // the temperature reset at nav happens in the step kernel
        sim.compute_energies();
        const int N = sim.N;
        double tempav = 0.0;  // not reset between trajectories (output_manager.cu:147-153)
        double rg = 0.0;      // likewise (output_manager.cu:196-204)
        for (int t = 0; t < sim.Ntr; ++t) {
            std::vector<float> T(N);
            sim.check(sim.eng->get_array(sim.jobs[t], 0, T.data()));
            for (int i = 0; i < N; ++i) tempav += T[i];
            tempav /= ((double)N * sim.nav);
            temperatures[t] = (float)tempav;
            const std::vector<float> c = sim.coords(t);
            int q = 0;
            for (size_t k = 0; k < sim.top.ni.size(); ++k) {
                const int i = sim.top.ni[k], j = sim.top.nj[k];
                const float rx = c[4 * i] - c[4 * j], ry = c[4 * i + 1] - c[4 * j + 1], rz = c[4 * i + 2] - c[4 * j + 2];
                const float dr = sqrtf(rx * rx + ry * ry + rz * rz);
                if (dr < R_limit_bond || std::fabs(sim.top.nr0[k] - dr) <= R_limit) q++;
            }
            nativeCounts[t] = q;
            if (computeRg) {
                for (int i = 0; i < N; ++i) {
                    float acc = 0.0f;
                    for (int j = 0; j < i; ++j) {
                        const float x = c[4 * j] - c[4 * i], y = c[4 * j + 1] - c[4 * i + 1], z = c[4 * j + 2] - c[4 * i + 2];
                        acc += x * x + y * y + z * z;
                    }
                    rg += acc;
                }
                rg = std::sqrt(rg / ((double)N * N));
                rgs[t] = (float)rg;
            }
        }
        const std::string runs =
            sim.Ntr == 1 ? "Run " + std::to_string(sim.firstrun) : std::to_string(sim.Ntr) + " runs";
        std::printf("Writing output at step %lld of %lld. %s on %d [%s].\n", sim.step, sim.numsteps, runs.c_str(),
                    sim.device, sim.device_name.c_str());
        title();
        for (int t = 0; t < sim.Ntr; ++t) {
            // the TEA column holds the float epsilon of the last update; the
            // exact (Cholesky) integrator leaves it unset, written here as 0
            float eps = 0.0f;
            if (sim.tea) {
                double e;
                float beta;
                sim.check(sim.eng->get_hi(sim.jobs[t], &e, &beta));
                eps = (float)e;
            }
            if (t < printRuns) screen(t, eps);
            if (t == printRuns) ldots();
            FILE* f = std::fopen(files[t].c_str(), "a");
            if (!f) throw ConfigError("Can't append to '" + files[t] + "'");
            std::fprintf(f, "%12lld\t%8.5f\t", sim.step, temperatures[t]);
            double total = 0.0;
            for (auto& pot : sim.potentials)
                for (int id = 0; id < pot->getEnergiesCount(); ++id) {
                    const float e = pot->getEnergy(t, id);
                    std::fprintf(f, "%8.5f\t", e);
                    total += e;
                }
            std::fprintf(f, "%8.5f\t", total);
            std::fprintf(f, "%8d\t", nativeCounts[t]);
            std::fprintf(f, "%8.5f", rgs[t]);
            if (sim.hi) std::fprintf(f, "\t%2.9f", eps);
            std::fprintf(f, "\n");
            std::fclose(f);
        }
    }
    void title() const {
        std::printf("%*s%*s%*s", width, "Run", width, "TimeStep", width, "Temperature");
        for (auto& pot : sim.potentials) {
            if (pot->getEnergiesCount() == 1) std::printf("%*s", width, pot->name.c_str());
            else
                for (int id = 0; id < pot->getEnergiesCount(); ++id)
                    std::printf("%*s%d", width - 1, pot->name.c_str(), id + 1);
        }
        std::printf("%*s%*s%*s", width, "Total", width, "Native#", width, "Rg");
        if (sim.hi) std::printf("%*s", width, "TEA-eps");
        std::printf("\n");
    }
    void screen(int t, float eps) const {
        std::printf("%*d%*lld%*f", width, t + sim.firstrun, width, sim.step, width, temperatures[t]);
        double total = 0.0;
        for (auto& pot : sim.potentials)
            for (int id = 0; id < pot->getEnergiesCount(); ++id) {
                const float e = pot->getEnergy(t, id);
                std::printf("%*f", width, e);
                total += e;
            }
        std::printf("%*f%*d%*f", width, total, width, nativeCounts[t], width, rgs[t]);
        if (sim.hi) std::printf("%*f", width, eps);
        std::printf("\n");
    }
    void ldots() const {
        std::printf("%*s%*s%*s", width, "...", width, "...", width, "...");
        for (auto& pot : sim.potentials)
            for (int id = 0; id < pot->getEnergiesCount(); ++id) std::printf("%*s", width, "...");
        std::printf("%*s%*s%*s", width, "...", width, "...", width, "...");
        if (sim.hi) std::printf("%*s", width, "...");
        std::printf("\n");
    }
};

struct DcdOutputManager : SOPUpdater {  // dcd_manager.cpp:35-107
    Sim& sim;
    int restartfreq;
    std::vector<std::string> dcd, restart;
    DcdOutputManager(Sim& s) : sim(s) {
        name = "DCD output";
        frequency = s.p.dcdfreq;
        restartfreq = (int)s.cfg.integer_or("restartfreq", 100000);
        const std::string dp = s.file_param("DCDfile", "<name>_<run>.dcd");
        const std::string rp = s.file_param("restartname", "restart.<name>_<run>.pdb");
        for (int t = 0; t < s.Ntr; ++t) {
            dcd.push_back(s.run_name(dp, t));
            restart.push_back(s.run_name(rp, t));
            DcdWriter::header(dcd.back(), s.N + (int)s.extra_atoms.size(), frequency, frequency, frequency, 0.0f);
        }
    }
    void update() override {
        if (sim.step % frequency == 0) {
            for (int t = 0; t < sim.Ntr; ++t) {
                const std::vector<float> c = sim.coords(t);
                const size_t m = sim.N + sim.extra_atoms.size();
                std::vector<float> X(m), Y(m), Z(m);
                for (int i = 0; i < sim.N; ++i) {
                    X[i] = c[4 * i];
                    Y[i] = c[4 * i + 1];
                    Z[i] = c[4 * i + 2];
                }
                for (size_t k = 0; k < sim.extra_atoms.size(); ++k) {
                    X[sim.N + k] = (float)sim.extra_atoms[k].x;
                    Y[sim.N + k] = (float)sim.extra_atoms[k].y;
                    Z[sim.N + k] = (float)sim.extra_atoms[k].z;
                }
                DcdWriter::frame(dcd[t], X, Y, Z);
            }
        }
        if (sim.step % restartfreq == 0) {
            for (int t = 0; t < sim.Ntr; ++t) {
                const std::vector<float> c = sim.coords(t);
                Topology copy = sim.top;
                for (int i = 0; i < sim.N; ++i) {
                    copy.atoms[i].x = c[4 * i];
                    copy.atoms[i].y = c[4 * i + 1];
                    copy.atoms[i].z = c[4 * i + 2];
                }
                write_pdb(restart[t], copy, sim.extra_atoms);
            }
        }
    }
};

struct TemperatureUpdater : SOPUpdater {  // langevin.cu:68-82
    Sim& sim;
    TemperatureUpdater(Sim& s) : sim(s) {
        name = "TemperatureUpdater";
        frequency = s.p.tempFreq;
    }
    void update() override {
        sim.check(sim.eng->stage_job_update(4));
        if (sim.step % frequency == 0)
            std::printf("Heating: %lld\t%f\n", sim.step,
                        sim.p.initialT + (float)(sim.step / frequency) * sim.p.deltaT);
    }
};

struct TeaUpdater : SOPUpdater {  // bdhitea.cu:126-176
    Sim& sim;
    TeaUpdater(Sim& s) : sim(s) {
        name = "BDHI-TEA updater";
        frequency = s.p.hi_epsilon_freq;
    }
    void update() override { sim.check(sim.eng->stage_tea_update()); }
};

struct StepIntegrator : SOPIntegrator {  // LangevinIntegrator / TeaIntegrator
    Sim& sim;
    StepIntegrator(Sim& s, const char* n) : sim(s) {
        name = n;
        h = s.p.timestep;
    }
    void integrate() override {
        // forces the host potentials added for this step
        for (int t = 0; t < sim.Ntr && !sim.host_forces_set.empty(); ++t) {
            if (!sim.host_forces_set[t]) continue;
            sim.check(sim.eng->upload_forces(sim.jobs[t], sim.host_forces[t].data()));
            std::fill(sim.host_forces[t].begin(), sim.host_forces[t].end(), 0.0f);
            sim.host_forces_set[t] = 0;
        }
        sim.check(sim.eng->stage_step(sim.substep++));
    }
};

// ---- configuration -> sop_params (the plugins' PARAMETER()s) ---------------
static void read_params(const Config& c, sop_params& p, int& nav) {
    std::memset(&p, 0, sizeof p);
    nav = (int)c.integer_or("nav", 1000);
    p.nav = nav;
    p.kspring_cov = (float)c.real_or("kspring_cov", 20.0);
    p.R_limit = (float)c.real_or("R_limit", 2.0);
    p.R_limit_bond = (float)c.real_or("R_limit_bond", 8.0);
    p.pairs_cutoff = (float)c.real_or("pairs_cutoff", 20.0);
    p.el = (float)c.real_or("el", 1.0);
    p.a = (float)c.real_or("a", 3.8);
    p.pairlist_cutoff = (float)c.real_or("pairlist_cutoff", 20.0);
    p.pairs_threshold = (float)c.real_or("pairs_threshold", 200.0);
    p.pairs_freq = (int)c.integer_or("pairs_freq", 1000);
    p.possiblepairs_freq = (int)c.integer_or("possiblepairs_freq", 100000);
    p.max_covalent = (int)c.integer_or("max_covalent", 8);
    p.max_native = (int)c.integer_or("max_native", 128);
    p.max_pairs = (int)c.integer_or("max_pairs", 512);
    p.max_possiblePairs = (int)c.integer_or("max_possiblePairs", 4096);
    p.timestep = (float)c.real("timestep");
    p.zeta = (float)c.real_or("zeta", 50.0);
    p.temperature = (float)c.real_or("temperature", 0.6);
    p.minimization = c.boolean_or("minimization", false);
    p.heating = c.boolean_or("heating", false);
    if (p.heating) {
        p.tempFreq = (int)c.integer("tempFreq");
        p.initialT = (float)c.real("initialT");
        p.deltaT = (float)c.real("deltaT");
    }
    p.outputtiming = (int)c.integer_or("outputtiming", 10000);
    p.dcdfreq = (int)c.integer_or("dcdfreq", 10000);
    if (c.boolean_or("pullingPlane", false))
        throw ConfigError("pullingPlane: the pulling-plane protocol is not part of this engine");
    p.pulling = c.boolean_or("pulling", false);
    if (p.pulling) {
        if (c.has("fixed_beads") || c.has("pulled_beads"))
            throw ConfigError("'fixed_beads' and 'pulled_beads' parameters are deprecated. Use 'fixed' and 'pulled' instead");
        p.deltax = (float)c.real_or("deltax", 0.0);
        p.k_trans = (float)c.real_or("k_trans", 0.05);
        p.fconst = (float)c.real_or("fconst", 0.0);
        if (p.deltax == 0.0f && p.fconst == 0.0f)
            throw ConfigError("ERROR: Either 'deltax' or 'fconst' parameter should be specified to initiate pulling");
        p.pullFreq = (int)c.integer_or("pullFreq", nav);
        const std::string dir = c.str_or("pullDirection", "endToEnd");
        if (dir == "vector") {
            p.pullDirectionVector = 1;
            const auto v = c.float3("pullVector");
            for (int d = 0; d < 3; ++d) p.pullVector[d] = v[d];
        } else if (dir != "endToEnd") {
            throw ConfigError("ERROR: 'pullDirection' parameter should be set to 'endToEnd' or 'vector'.");
        }
        p.fixedEnd = (int)c.integer("fixedEnd");
        p.pulledEnd = (int)c.integer("pulledEnd");
    }
    p.indentation = c.boolean_or("indentation", false);
    if (p.indentation) {
        if (c.boolean_or("indentationDiscreteSurf", false))
            throw ConfigError("indentationDiscreteSurf: the discrete mica surface is not part of this engine");
        const auto chip = c.float3("indentationChip");
        const auto tip = c.float3_or("indentationTip", chip);
        const auto dir = c.float3("indentationDirection");
        const auto r0 = c.float3("indentationSurfaceR0");
        const auto nrm = c.float3("indentationSurfaceN");
        for (int d = 0; d < 3; ++d) {
            p.indChip[d] = chip[d];
            p.indTip[d] = tip[d];
            p.indDirection[d] = dir[d];
            p.indSurfR0[d] = r0[d];
            p.indSurfN[d] = nrm[d];
        }
        p.indTipR = (float)c.real("indentationTipR");
        p.indKs = (float)c.real("indentationKs");
        p.indDeltaX = (float)c.real("indentationDeltaX");
        p.indMoveSurface = c.boolean_or("indentationMoveSurface", false);
        p.indFixTrans = c.boolean_or("indentationFixTrans", false);
        const float sigma = (float)c.real_or("indentationSigma", 1.0);
        const float el = (float)c.real_or("indentationEl", 1.0);
        p.indTipSigma = (float)c.real_or("indentationTipSigma", sigma);
        p.indSurfSigma = (float)c.real_or("indentationSurfSigma", sigma);
        p.indTipEl = (float)c.real_or("indentationTipEl", el);
        p.indSurfEl = (float)c.real_or("indentationSurfEl", el);
        p.indTipA = (float)c.real_or("indentationTipA", 0.0);
        p.indTipB = (float)c.real_or("indentationTipB", 1.0);
        p.indSurfA = (float)c.real_or("indentationSurfA", 0.0);
        p.indSurfB = (float)c.real_or("indentationSurfB", 1.0);
        p.indTipZeta = (float)c.real_or("indentationTipZeta", 5000.0);
        p.indOutputFreq = (int)c.integer_or("indentationOutputFreq", 1000);
        p.indRetractionStep = c.integer_or("indentationRetractionStep", -1);
    }
    p.hi_on = c.boolean_or("hi_on", false);
    if (p.hi_on && p.minimization)
        throw ConfigError("Using minimization with BDHI integrator is not supported and will never be.");
    if (p.hi_on) {
        p.hi_exact = c.boolean_or("hi_exact", false);
        p.hi_a = (float)c.real_or("hi_a", 1.8);
        if (!p.hi_exact) {
            p.hi_capricious = c.boolean_or("hi_capricious", true);
            p.hi_epsilon_freq = (int)c.integer("hi_epsilon_freq");
            p.hi_epsmax = (float)c.real_or("hi_epsmax", 999.0);
            p.hi_unlisted = c.boolean_or("hi_unlisted", true);
        }
        p.hi_chol_freq = (int)c.integer_or("hi_chol_freq", 1);
        p.hi_faithful_blocks = c.boolean_or("hi_faithful_blocks", true);
    }
}

// initGPU + initFF (src/gsop.cu:45-106, src/sop-gpu.cpp:18-45): parameters,
// topology, coordinates and seeds of every trajectory, the engine, and the
// built-in plugins in initFF's order.  The trajectories are admitted on the
// GPU by run_sim(), once the host has added its plugins.
static void create_sim(Sim& sim, const std::string& path,
                       const std::vector<std::pair<std::string, std::string>>& overrides) {
    sim.t_start = std::time(nullptr);
    std::printf("===========================\nSOP-GPU (MI355X engine) %s\n===========================\n", sop_version());
    sim.cfg.load(path);
    for (auto& kv : overrides) sim.cfg.set(kv.first, kv.second);
    Config& c = sim.cfg;
    if (c.has("stage"))
        throw ConfigError("Parameter `stage` is no longer supported. Please enable all necessary protocols through "
                          "individual parameters.");
    std::printf("Initializing simulations for '%s'\n", c.str_or("name", "unnamed").c_str());
    const int run = (int)c.integer_or("run", -1);
    if (run == -1) {
        sim.Ntr = (int)c.integer("runnum");
        sim.firstrun = (int)c.integer("firstrun");
    } else {
        sim.Ntr = 1;
        sim.firstrun = run;
    }
    sim.numsteps = c.integer("numsteps");
    read_params(c, sim.p, sim.nav);
    const int device = (int)c.integer_or("device", 0);
    sim.device = device;
    sim.top = read_topology(c.str("topology"));
    sim.N = (int)sim.top.atoms.size();
    const int N = sim.N;
    // coordinates of every trajectory (gsop.cu:185-199, sop-gpu.cpp:41-45)
    sim.xyz.assign(sim.Ntr, {});
    for (int t = 0; t < sim.Ntr; ++t) {
        const std::string cf = string_replace(c.str("coordinates"), "<run>", std::to_string(t + sim.firstrun));
        read_coordinates(cf, sim.top);
        sim.xyz[t].resize(4 * (size_t)N);
        for (int i = 0; i < N; ++i) {
            sim.xyz[t][4 * i] = (float)sim.top.atoms[i].x;
            sim.xyz[t][4 * i + 1] = (float)sim.top.atoms[i].y;
            sim.xyz[t][4 * i + 2] = (float)sim.top.atoms[i].z;
            sim.xyz[t][4 * i + 3] = 0.0f;
        }
    }
    // seed + firstrun, one seed table over the concatenated beads (langevin.cu:23-24)
    const int seed = (int)(c.has("seed") ? c.integer("seed") : (long long)std::time(nullptr)) + sim.firstrun;
    sim.p.seed = seed;
    sim.seeds.resize(4 * (size_t)N * sim.Ntr);
    generate_seeds(seed, N * sim.Ntr, sim.seeds.data());
    if (sim.p.pulling) {
        sim.fixed = c.int_array("fixed");
        sim.pulled = c.int_array("pulled");
        std::printf("%zu resid(s) fixed, %zu pulled.\n", sim.fixed.size(), sim.pulled.size());
        // the pulled/fixed flags overwrite the B-factor column (pulling.cu:123-125)
        for (int i : sim.pulled)
            if (i > N || i < 0)
                throw ConfigError("Invalid pulling bead #" + std::to_string(i) + "; should be >=0 and <" + std::to_string(N));
        for (int i : sim.fixed)
            if (i > N || i < 0)
                throw ConfigError("Invalid fixed bead #" + std::to_string(i) + "; should be >=0 and <" + std::to_string(N));
        for (auto& a : sim.top.atoms) a.beta = 0.0;
        for (int i : sim.pulled)
            if (i >= 0 && i < N) sim.top.atoms[i].beta = 2.0;
        for (int i : sim.fixed)
            if (i >= 0 && i < N) sim.top.atoms[i].beta = 1.0;
    }
    if (sim.p.indentation && sim.Ntr > 1)
        throw ConfigError("Many-runs-per-GPU mode is not supported by indentation potential.");
    size_t free_b = 0, total_b = 0;
    if (hipSetDevice(device) != hipSuccess || hipMemGetInfo(&free_b, &total_b) != hipSuccess)
        throw ConfigError("Can't use HIP device " + std::to_string(device));
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) sim.device_name = prop.name;
    std::printf("Using device %d: \"%s\"\n", device, sim.device_name.c_str());
    sim.eng = std::make_unique<Engine>(device, (size_t)(0.9 * (double)free_b));
    sim.eng->set_tuning(0, 0);
    if (c.has("hi_chol_mode")) sim.eng->set_exact_mode((int)c.integer("hi_chol_mode"));
    // plugins in initFF's order (gsop.cu:80-101)
    sim.hi = sim.p.hi_on;
    sim.tea = sim.p.hi_on && !sim.p.hi_exact;
    sim.potentials.emplace_back(new TermPotential(sim, "Covalent", 0));
    sim.potentials.emplace_back(new TermPotential(sim, "Native", 1));
    sim.potentials.emplace_back(new TermPotential(sim, "Long Range", 2));
    if (sim.p.indentation) {
        const bool show = c.boolean_or("indentationShowTipSurf", false);
        indentation_representation(sim);
        sim.updaters.emplace_back(new IndentationTipUpdater(sim));
        if (show) sim.updaters.emplace_back(new IndentationAminoUpdater(sim));
        else sim.extra_atoms.clear();
        sim.potentials.emplace_back(new KernelPotential("Indentation"));
    }
    if (sim.p.pulling) {
        sim.potentials.emplace_back(new KernelPotential("Pulling"));
        if (sim.p.deltax != 0.0f) sim.updaters.emplace_back(new PullingUpdater(sim));
    }
    sim.updaters.emplace_back(new PossiblepairUpdater(sim));
    sim.updaters.emplace_back(new OutputManager(sim));
    sim.updaters.emplace_back(new PairlistUpdater(sim));
    sim.updaters.emplace_back(new DcdOutputManager(sim));
    if (sim.p.heating) sim.updaters.emplace_back(new TemperatureUpdater(sim));
    if (sim.tea) sim.updaters.emplace_back(new TeaUpdater(sim));
    sim.integrator.reset(new StepIntegrator(sim, sim.hi ? "BDHI-TEA" : "Langevin"));
    sim.builtin_updaters = sim.updaters.size();
}

// The trajectories on the GPU, runGPU (gsop.cu:111-163) and the final
// coordinates (sop-gpu.cpp:47-65).
static void run_sim(Sim& sim) {
    const Config& c = sim.cfg;
    const int N = sim.N;
    // host plugins: their updaters' frequencies bound the stride, their
    // potentials' forces reach the step through the job's force array
    int host_min = 0;
    for (size_t k = sim.builtin_updaters; k < sim.updaters.size(); ++k)
        host_min = host_min ? std::min(host_min, sim.updaters[k]->frequency) : sim.updaters[k]->frequency;
    sim.p.updater_freq_min = host_min;
    bool host_pot = false;
    for (auto& pot : sim.potentials)
        if (dynamic_cast<HostPotential*>(pot.get())) host_pot = true;
    sim.p.host_forces = host_pot ? 1 : 0;
    if (host_pot) {
        sim.host_forces.assign(sim.Ntr, std::vector<float>(4 * (size_t)N, 0.0f));
        sim.host_forces_set.assign(sim.Ntr, 0);
    }
    sim.running = true;
    for (int t = 0; t < sim.Ntr; ++t) {
        sop_system s{};
        s.n = N;
        s.coords4 = sim.xyz[t].data();
        s.nbonds = (int)sim.top.bi.size();
        s.bond_i = sim.top.bi.data();
        s.bond_j = sim.top.bj.data();
        s.bond_r0 = sim.top.br0.data();
        s.nnative = (int)sim.top.ni.size();
        s.nat_i = sim.top.ni.data();
        s.nat_j = sim.top.nj.data();
        s.nat_r0 = sim.top.nr0.data();
        s.nat_eh = sim.top.neh.data();
        s.npairs = (int)sim.top.pi.size();
        s.pair_i = sim.top.pi.data();
        s.pair_j = sim.top.pj.data();
        s.nfixed = (int)sim.fixed.size();
        s.fixed = sim.fixed.data();
        s.npulled = (int)sim.pulled.size();
        s.pulled = sim.pulled.data();
        s.rng_state4 = sim.seeds.data() + 4 * (size_t)N * t;
        int id = 0;
        sim.check(sim.eng->add_job(sim.p, s, sim.numsteps, &id));
        sim.jobs.push_back(id);
    }

    const std::string ref = sim.file_param("reffilename", "<name>.ref.pdb");
    std::printf("Saving reference PDB: '%s'.\n", ref.c_str());
    write_pdb(ref, sim.top, sim.extra_atoms);  // holds the last trajectory's coordinates, as in the reference
    std::printf("System initialization took %.f second(s).\n", difftime(std::time(nullptr), sim.t_start));
    std::printf("Starting %d runs (%lld steps).\n", sim.Ntr, sim.numsteps);

    // runGPU (gsop.cu:111-163)
    int stride = sim.nav;
    for (auto& u : sim.updaters)
        if (u->frequency < stride) stride = u->frequency;
    // the Verlet list is built once before the loop, from the possible list
    // of the topology's [ pairs ], so the step-0 energies use it (gsop.cu:126)
    for (size_t k = 0; k < sim.builtin_updaters; ++k)
        if (sim.updaters[k]->name == "Pairlist") sim.updaters[k]->update();
    for (long long b = sim.step / stride; b <= sim.numsteps / stride; ++b) {
        sim.substep = 0;
        for (auto& u : sim.updaters) u->update();
        for (int j = 0; j < stride; ++j) {
            for (auto& pot : sim.potentials) pot->compute();
            sim.integrator->integrate();
        }
        sim.check(sim.eng->stage_advance(stride));
        sim.step += stride;
        sim.substep = 0;
    }
    const std::string fin = sim.file_param("finalcoord", "<name>_<run>_final.pdb");
    for (int t = 0; t < sim.Ntr; ++t) {
        const std::vector<float> xf = sim.coords(t);
        Topology copy = sim.top;
        for (int i = 0; i < N; ++i) {
            copy.atoms[i].x = xf[4 * i];
            copy.atoms[i].y = xf[4 * i + 1];
            copy.atoms[i].z = xf[4 * i + 2];
        }
        write_pdb(sim.run_name(fin, t), copy, sim.extra_atoms);
    }
    long long st = 0, end = 0;
    int ovf = 0;
    for (int id : sim.jobs) {
        sim.check(sim.eng->job_status(id, &st, &end, &ovf));
        if (ovf & 3) throw ConfigError("a pairlist outgrew its capacity (max_pairs / max_possiblePairs)");
        if (ovf & 12) throw ConfigError("HI tensor check failed (hi_capricious)");
    }
    (void)c;
}

}  // namespace sop

struct sop_sim {
    std::unique_ptr<sop::Sim> sim;
};

namespace {
std::vector<std::pair<std::string, std::string>> override_list(int n, const char* const* names,
                                                               const char* const* values) {
    std::vector<std::pair<std::string, std::string>> ov;
    for (int k = 0; k < n; ++k) ov.push_back({names[k], values[k]});
    return ov;
}

template <typename F>
int guarded(F&& f) {
    try {
        return f();
    } catch (const sop::ConfigError& e) {
        return sop::fail(SOP_E_CONFIG, e.what());
    } catch (const std::exception& e) {
        return sop::fail(SOP_E_CONFIG, e.what());
    }
}
}  // namespace

extern "C" {

int sop_run_config(const char* path, int noverride, const char* const* names, const char* const* values) {
    if (!path || (noverride > 0 && (!names || !values))) return sop::fail(SOP_E_ARG, "null path or overrides");
    return guarded([&] {
        sop::Sim sim;
        sop::create_sim(sim, path, override_list(noverride, names, values));
        sop::run_sim(sim);
        return 0;
    });
}

int sop_sim_create(const char* path, int noverride, const char* const* names, const char* const* values,
                   sop_sim** out) {
    if (!out) return sop::fail(SOP_E_ARG, "null output");
    *out = nullptr;
    if (!path || (noverride > 0 && (!names || !values))) return sop::fail(SOP_E_ARG, "null path or overrides");
    auto* h = new sop_sim;
    h->sim = std::make_unique<sop::Sim>();
    h->sim->handle = h;
    const int rc = guarded([&] {
        sop::create_sim(*h->sim, path, override_list(noverride, names, values));
        return 0;
    });
    if (rc) {
        delete h;
        return rc;
    }
    *out = h;
    return 0;
}

int sop_sim_run(sop_sim* s) {
    if (!s || !s->sim) return sop::fail(SOP_E_ARG, "null simulation");
    if (s->sim->ran) return sop::fail(SOP_E_STATE, "the simulation has already run");
    s->sim->ran = true;
    return guarded([&] {
        sop::run_sim(*s->sim);
        return 0;
    });
}

void sop_sim_destroy(sop_sim* s) { delete s; }

void sop_sim_fail(const char* message) { sop::g_plugin_error = message ? message : ""; }

int sop_sim_add_updater(sop_sim* s, const char* name, int frequency, sop_plugin_fn update, void* user,
                        sop_release_fn release) {
    int rc = 0;
    if (!s || !s->sim || !name || !update) rc = sop::fail(SOP_E_ARG, "null argument");
    else if (frequency <= 0) rc = sop::fail(SOP_E_ARG, "updater frequency must be positive");
    else if (s->sim->ran) rc = sop::fail(SOP_E_STATE, "plugins are registered before the run");
    if (rc) {
        if (release) release(user);
        return rc;
    }
    s->sim->updaters.emplace_back(new sop::HostUpdater(s, name, frequency, update, user, release));
    return 0;
}

int sop_sim_add_potential(sop_sim* s, const char* name, int energies_count, sop_plugin_fn compute,
                          sop_energy_fn energy, void* user, sop_release_fn release) {
    int rc = 0;
    if (!s || !s->sim || !name) rc = sop::fail(SOP_E_ARG, "null argument");
    else if (energies_count < 0 || (energies_count > 0 && !energy)) rc = sop::fail(SOP_E_ARG, "bad energy terms");
    else if (s->sim->ran) rc = sop::fail(SOP_E_STATE, "plugins are registered before the run");
    if (rc) {
        if (release) release(user);
        return rc;
    }
    s->sim->potentials.emplace_back(new sop::HostPotential(s, name, energies_count, compute, energy, user, release));
    return 0;
}

int sop_sim_potential_count(const sop_sim* s, int* count) {
    if (!s || !s->sim || !count) return sop::fail(SOP_E_ARG, "null argument");
    *count = (int)s->sim->potentials.size();
    return 0;
}

int sop_sim_updater_count(const sop_sim* s, int* count) {
    if (!s || !s->sim || !count) return sop::fail(SOP_E_ARG, "null argument");
    *count = (int)s->sim->updaters.size();
    return 0;
}

const char* sop_sim_potential_name(const sop_sim* s, int index) {
    if (!s || !s->sim || index < 0 || index >= (int)s->sim->potentials.size()) return nullptr;
    return s->sim->potentials[index]->name.c_str();
}

const char* sop_sim_updater_name(const sop_sim* s, int index) {
    if (!s || !s->sim || index < 0 || index >= (int)s->sim->updaters.size()) return nullptr;
    return s->sim->updaters[index]->name.c_str();
}

int sop_sim_updater_frequency(const sop_sim* s, int index, int* frequency) {
    if (!s || !s->sim || !frequency || index < 0 || index >= (int)s->sim->updaters.size())
        return sop::fail(SOP_E_ARG, "bad argument");
    *frequency = s->sim->updaters[index]->frequency;
    return 0;
}

int sop_sim_potential_by_name(const sop_sim* s, const char* name, int* index) {
    if (!s || !s->sim || !name || !index) return sop::fail(SOP_E_ARG, "null argument");
    for (size_t k = 0; k < s->sim->potentials.size(); ++k)
        if (s->sim->potentials[k]->name == name) {
            *index = (int)k;
            return 0;
        }
    return sop::fail(SOP_E_ARG, std::string("Unable to find requested potential '") + name + "'");
}

int sop_sim_updater_by_name(const sop_sim* s, const char* name, int* index) {
    if (!s || !s->sim || !name || !index) return sop::fail(SOP_E_ARG, "null argument");
    for (size_t k = 0; k < s->sim->updaters.size(); ++k)
        if (s->sim->updaters[k]->name == name) {
            *index = (int)k;
            return 0;
        }
    return sop::fail(SOP_E_ARG, std::string("Unable to find requested updater '") + name + "'");
}

const char* sop_sim_integrator_name(const sop_sim* s) {
    if (!s || !s->sim || !s->sim->integrator) return nullptr;
    return s->sim->integrator->name.c_str();
}

int sop_sim_get_energy(sop_sim* s, int index, int traj, int id, float* energy) {
    if (!s || !s->sim || !energy) return sop::fail(SOP_E_ARG, "null argument");
    sop::Sim& sim = *s->sim;
    if (index < 0 || index >= (int)sim.potentials.size() || traj < 0 || traj >= sim.Ntr || id < 0 ||
        id >= sim.potentials[index]->getEnergiesCount())
        return sop::fail(SOP_E_ARG, "no such energy term");
    if (sim.energies.empty()) return sop::fail(SOP_E_STATE, "no energies computed yet");
    *energy = sim.potentials[index]->getEnergy(traj, id);
    return 0;
}

int sop_sim_state(const sop_sim* s, long long* step, int* n, int* ntr) {
    if (!s || !s->sim) return sop::fail(SOP_E_ARG, "null simulation");
    if (step) *step = s->sim->step + s->sim->substep;
    if (n) *n = s->sim->N;
    if (ntr) *ntr = s->sim->Ntr;
    return 0;
}

int sop_sim_get_coords(sop_sim* s, int traj, float* out4) {
    if (!s || !s->sim || !out4) return sop::fail(SOP_E_ARG, "null argument");
    sop::Sim& sim = *s->sim;
    if (traj < 0 || traj >= sim.Ntr) return sop::fail(SOP_E_ARG, "no such trajectory");
    if (sim.jobs.empty()) {  // not admitted yet: the initial coordinates
        std::memcpy(out4, sim.xyz[traj].data(), sim.xyz[traj].size() * sizeof(float));
        return 0;
    }
    return sim.eng->get_coords_at(sim.jobs[traj], sim.step + sim.substep, out4);
}

int sop_sim_add_forces(sop_sim* s, int traj, const float* f4) {
    if (!s || !s->sim || !f4) return sop::fail(SOP_E_ARG, "null argument");
    sop::Sim& sim = *s->sim;
    if (traj < 0 || traj >= sim.Ntr) return sop::fail(SOP_E_ARG, "no such trajectory");
    if (!sim.running || sim.host_forces.empty())
        return sop::fail(SOP_E_STATE, "forces are added by a registered potential's compute() during the run");
    std::vector<float>& acc = sim.host_forces[traj];
    for (size_t i = 0; i < acc.size(); ++i) acc[i] += (i & 3) == 3 ? 0.0f : f4[i];
    sim.host_forces_set[traj] = 1;
    return 0;
}

}  // extern "C"
