// Ensembles of independent contexts stepped from captured HIP graphs.
//
// Each member keeps its own context (grid, operators, state) and stream.  One ensemble
// step is one graph: the master stream forks to every member's stream, each member
// enqueues its whole time step (every kernel of navierstokes_update, norm and the time
// update), and the members join back; the members' kernels are independent branches
// of the graph and run concurrently, so small grids of uneven sizes fill the GPU
// together with one launch per step.  The history arrays rotate by slot index each
// step, so the kernels' arguments repeat only after the rotation period (60 steps);
// the ensemble captures one graph per phase of that period and launches the graph of
// the current phase, advancing every member's host-side bookkeeping (step counter,
// slot rotation, time) as the step it stands for would.  Captured steps need steady
// EX3/BDF3 coefficients, so the members run their start-up steps before capture.

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "lsb_internal.h"

struct lsb_ensemble {
    std::vector<lsb_ctx*> m;
    hipStream_t master = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    std::vector<hipGraphExec_t> exec;
    int phase = 0;
    int device = 0;
};

using namespace lsb;

namespace {

void release(lsb_ensemble* e) {
    for (hipGraphExec_t x : e->exec) (void)hipGraphExecDestroy(x);
    e->exec.clear();
    for (hipEvent_t x : e->join) (void)hipEventDestroy(x);
    e->join.clear();
    if (e->fork) (void)hipEventDestroy(e->fork);
    e->fork = nullptr;
    if (e->master) (void)hipStreamDestroy(e->master);
    e->master = nullptr;
    for (lsb_ctx* c : e->m) {
        c->graphs_live = false;
        c->stage_locked = false;
        c->capturing = false;
        c->stage_in_graph.assign(c->stage_in_graph.size(), 0);
    }
}

// master waits for the members' earlier work; members' later work waits for the master
int fork_join_host(lsb_ensemble* e, bool into_master) {
    if (into_master) {
        for (size_t i = 0; i < e->m.size(); ++i) {
            LSB_HIP(hipEventRecord(e->join[i], e->m[i]->stream));
            LSB_HIP(hipStreamWaitEvent(e->master, e->join[i], 0));
        }
    } else {
        LSB_HIP(hipEventRecord(e->fork, e->master));
        for (size_t i = 0; i < e->m.size(); ++i) LSB_HIP(hipStreamWaitEvent(e->m[i]->stream, e->fork, 0));
    }
    return LSB_OK;
}

int capture_phase(lsb_ensemble* e, hipGraphExec_t* out) {
    LSB_HIP(hipStreamBeginCapture(e->master, hipStreamCaptureModeThreadLocal));
    int rc = LSB_OK;
    hipError_t he = hipEventRecord(e->fork, e->master);
    for (size_t i = 0; i < e->m.size() && rc == LSB_OK && he == hipSuccess; ++i) {
        lsb_ctx* c = e->m[i];
        he = hipStreamWaitEvent(c->stream, e->fork, 0);
        if (he == hipSuccess) rc = step_enqueue(c);
        if (rc == LSB_OK && he == hipSuccess) he = hipEventRecord(e->join[i], c->stream);
        if (rc == LSB_OK && he == hipSuccess) he = hipStreamWaitEvent(e->master, e->join[i], 0);
    }
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(e->master, &g);
    if (rc != LSB_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (he != hipSuccess || ee != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        set_error(std::string("ensemble capture: ") + hipGetErrorString(he != hipSuccess ? he : ee));
        return LSB_E_HIP;
    }
    const hipError_t ie = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) {
        set_error(std::string("ensemble graph instantiation: ") + hipGetErrorString(ie));
        return LSB_E_HIP;
    }
    return LSB_OK;
}

}  // namespace

extern "C" int lsb_ensemble_create(lsb_ctx** members, int n, lsb_ensemble** out) {
    if (!members || n < 1 || !out) {
        set_error("ensemble_create: need n >= 1 member contexts");
        return LSB_E_ARG;
    }
    *out = nullptr;
    const int dev = members[0]->device;
    int period = 0;
    for (int i = 0; i < n; ++i) {
        lsb_ctx* c = members[i];
        std::string why;
        if (!c || c->device != dev || c->graphs_live) {
            set_error("ensemble_create: members must be distinct contexts on one device, not in another ensemble");
            return LSB_E_ARG;
        }
        if (!step_capturable(c, why)) {
            set_error("ensemble_create: member " + std::to_string(i) + " cannot be captured: " + why);
            return LSB_E_STATE;
        }
        for (int j = 0; j < i; ++j)
            if (members[j] == c) {
                set_error("ensemble_create: a context appears twice");
                return LSB_E_ARG;
            }
        const int p = step_rotation_period(c);
        if (p <= 0 || (period && p != period)) {
            set_error("ensemble_create: members' slot rotations differ");
            return LSB_E_STATE;
        }
        period = p;
    }
    LSB_HIP(hipSetDevice(dev));
    lsb_ensemble* e = new lsb_ensemble();
    e->m.assign(members, members + n);
    e->device = dev;
    int rc = LSB_OK;
    if (hipStreamCreate(&e->master) != hipSuccess || hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) != hipSuccess)
        rc = LSB_E_HIP;
    e->join.assign(n, nullptr);
    for (int i = 0; i < n && rc == LSB_OK; ++i)
        if (hipEventCreateWithFlags(&e->join[i], hipEventDisableTiming) != hipSuccess) rc = LSB_E_HIP;
    if (rc != LSB_OK) {
        set_error("ensemble_create: stream / event creation failed");
        release(e);
        delete e;
        return rc;
    }
    // capture one graph per phase of the rotation; capturing advances the host-side
    // bookkeeping by one step per phase, which a full period returns to its start
    // except for the counters, restored afterwards
    struct Saved { int istep, isum; double t, dt2, dt3; };
    std::vector<Saved> sv(n);
    for (int i = 0; i < n; ++i) {
        lsb_ctx* c = e->m[i];
        sv[i] = Saved{c->istep, c->isum, c->t, c->dt2, c->dt3};
        c->stage_locked = true;
        c->capturing = true;
    }
    for (int ph = 0; ph < period && rc == LSB_OK; ++ph) {
        hipGraphExec_t x = nullptr;
        rc = capture_phase(e, &x);
        if (rc == LSB_OK) e->exec.push_back(x);
    }
    for (int i = 0; i < n; ++i) {
        lsb_ctx* c = e->m[i];
        c->istep = sv[i].istep;
        c->isum = sv[i].isum;
        c->t = sv[i].t;
        c->dt2 = sv[i].dt2;
        c->dt3 = sv[i].dt3;
        c->capturing = false;
        c->graphs_live = (rc == LSB_OK);
    }
    if (rc != LSB_OK) {
        // a capture that failed partway leaves the members' slot bookkeeping mid-period:
        // their state must be set again
        for (lsb_ctx* c : e->m) c->state_ready = false;
        release(e);
        delete e;
        return rc;
    }
    *out = e;
    return LSB_OK;
}

extern "C" int lsb_ensemble_timestep(lsb_ensemble* e, int nsteps) {
    if (!e || nsteps < 0) {
        set_error("ensemble_timestep: null ensemble or nsteps < 0");
        return LSB_E_ARG;
    }
    LSB_HIP(hipSetDevice(e->device));
    LSB_TRY(fork_join_host(e, true));
    const int period = (int)e->exec.size();
    for (int k = 0; k < nsteps; ++k) {
        LSB_HIP(hipGraphLaunch(e->exec[e->phase], e->master));
        for (lsb_ctx* c : e->m) step_host_advance(c);
        e->phase = (e->phase + 1) % period;
    }
    return fork_join_host(e, false);
}

extern "C" int lsb_ensemble_synchronize(lsb_ensemble* e) {
    if (!e) {
        set_error("ensemble_synchronize: null ensemble");
        return LSB_E_ARG;
    }
    LSB_HIP(hipStreamSynchronize(e->master));
    for (lsb_ctx* c : e->m) LSB_HIP(hipStreamSynchronize(c->stream));
    return LSB_OK;
}

extern "C" int lsb_ensemble_info(const lsb_ensemble* e, int* nmembers, int* ngraphs, int* phase) {
    if (!e) {
        set_error("ensemble_info: null ensemble");
        return LSB_E_ARG;
    }
    if (nmembers) *nmembers = (int)e->m.size();
    if (ngraphs) *ngraphs = (int)e->exec.size();
    if (phase) *phase = e->phase;
    return LSB_OK;
}

extern "C" int lsb_ensemble_destroy(lsb_ensemble* e) {
    if (!e) return LSB_OK;
    (void)hipSetDevice(e->device);
    if (e->master) (void)hipStreamSynchronize(e->master);
    release(e);
    delete e;
    return LSB_OK;
}
