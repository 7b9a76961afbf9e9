/*
 * sop_plugin.h -- the reference's plugin interfaces on the C ABI.
 *
 * The reference extends its engine with C++ plugins: a potential adds
 * forces every step, an updater runs every `stride` steps and checks its
 * own frequency, one integrator advances the coordinates
 * (SOPPotential / SOPUpdater / SOPIntegrator, src/gsop.h:69-105).  They
 * live in fixed registries filled in initFF's order (src/gsop.cu:31-34,
 * 80-101) and are found with potentialByName / updaterByName
 * (src/gsop.cu:237-253).
 *
 * This header keeps those classes.  The built-in potentials and updaters of
 * the sop-gpu program are objects of them inside the library; a host adds
 * its own by subclassing and registering between sop_sim_create() and
 * sop_sim_run().  Registration goes through the C ABI of sopgpu.h with
 * trampolines compiled into the caller (this header is header-only), so no
 * C++ object layout crosses the library boundary.  A registered object is
 * owned by the simulation and deleted by sop_sim_destroy().
 *
 * Order: the registries hold the built-ins in initFF's order, then the
 * host's plugins in registration order.  Every `stride` steps each updater's
 * update() runs in that order (stride = min(nav, every updater's
 * frequency), so a host updater's frequency shortens the stride as any
 * updater's does in the reference); then, for each of the stride steps,
 * every potential's compute() and the integrator.  An update() or compute()
 * that throws ends the run with SOP_E_CONFIG and the exception's text.
 *
 * Threading: one simulation per host thread; the callbacks run on the
 * thread that called sop_sim_run().
 */
#ifndef SOP_PLUGIN_H
#define SOP_PLUGIN_H

#include <exception>
#include <string>

#include "sopgpu.h"

namespace sop {

class SOPPotential {  // src/gsop.h:69-85
   public:
    virtual ~SOPPotential() {}
    std::string name;
    /* Add this potential's forces for the current step: sop_sim_add_forces(). */
    virtual void compute() = 0;
    /* Energy terms this potential reports in the energy .dat columns. */
    virtual int getEnergiesCount() const { return 0; }
    virtual float getEnergy(int traj, int id) { return 0.0f; }
};

class SOPUpdater {  // src/gsop.h:87-94
   public:
    virtual ~SOPUpdater() {}
    std::string name;
    int frequency = 1;
    /* Runs every stride steps; checks `frequency` itself (sop_sim_state). */
    virtual void update() = 0;
};

class SOPIntegrator {  // src/gsop.h:96-103 (the program's integrator; read-only here)
   public:
    virtual ~SOPIntegrator() {}
    std::string name;
    float h = 0.0f;
    virtual void integrate() = 0;
};

namespace plugin_detail {
inline int run_update(sop_sim*, void* user) {
    try {
        static_cast<SOPUpdater*>(user)->update();
        return 0;
    } catch (const std::exception& e) {
        sop_sim_fail(e.what());
        return -1;
    }
}
inline int run_compute(sop_sim*, void* user) {
    try {
        static_cast<SOPPotential*>(user)->compute();
        return 0;
    } catch (const std::exception& e) {
        sop_sim_fail(e.what());
        return -1;
    }
}
inline float energy(sop_sim*, int traj, int id, void* user) {
    return static_cast<SOPPotential*>(user)->getEnergy(traj, id);
}
inline void release_updater(void* user) { delete static_cast<SOPUpdater*>(user); }
inline void release_potential(void* user) { delete static_cast<SOPPotential*>(user); }
}  // namespace plugin_detail

/* Appends u to the updater registry; the simulation owns it from here on
 * (also when the call fails). */
inline int addUpdater(sop_sim* sim, SOPUpdater* u) {
    const int rc = sop_sim_add_updater(sim, u->name.c_str(), u->frequency, plugin_detail::run_update, u,
                                       plugin_detail::release_updater);
    return rc;
}

/* Appends p to the potential registry; the simulation owns it from here on. */
inline int addPotential(sop_sim* sim, SOPPotential* p) {
    return sop_sim_add_potential(sim, p->name.c_str(), p->getEnergiesCount(), plugin_detail::run_compute,
                                 plugin_detail::energy, p, plugin_detail::release_potential);
}

/* potentialByName / updaterByName (src/gsop.cu:237-253): the registry
 * index, or -1 with sop_last_error() set ("Unable to find requested ..."). */
inline int potentialByName(sop_sim* sim, const char* name) {
    int idx = -1;
    return sop_sim_potential_by_name(sim, name, &idx) == 0 ? idx : -1;
}
inline int updaterByName(sop_sim* sim, const char* name) {
    int idx = -1;
    return sop_sim_updater_by_name(sim, name, &idx) == 0 ? idx : -1;
}

}  // namespace sop

#endif
