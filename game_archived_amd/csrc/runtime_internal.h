/*
 * runtime_internal.h -- the device-resident model of the runtime
 * (rkhevi_runtime.hip), shared with the adjoint window (rkhevi_window.hip).
 * Not part of the C ABI.
 */
#ifndef GAME_RUNTIME_INTERNAL_H
#define GAME_RUNTIME_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string>
#include <vector>

#include "../../include/game_rkhevi.h"
#include "rkhevi_dev.h"
#include "rkhevi_kernels.h"
#include "rkhevi_adjoint.h"
#include "step_adjoint.h"

namespace rt {

extern char g_err[512];

/* integrals formed by one game_gpu_integrals call: the densities (at most six
   constituents), rhotheta_v and three energies */
const int INTEGRALS_MAX = 16;

#define HIP_OK(call)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            snprintf(g_err, sizeof g_err, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                     __FILE__, __LINE__);                                                   \
            return GAME_ERR_HIP;                                                            \
        }                                                                                   \
    } while (0)

struct Stat {
    std::string name;
    long long launches = 0;
    double ms = 0.0;
};

struct EventTimer : KernelTimer {
    std::vector<Stat> stats;
    struct Pending { int idx; hipEvent_t a, b; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    int cur = -1;
    hipEvent_t cur_a = nullptr;

    hipEvent_t get() {
        if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
        hipEvent_t e;
        (void)hipEventCreate(&e);
        return e;
    }
    int index(const char *name) {
        for (size_t k = 0; k < stats.size(); ++k) if (stats[k].name == name) return (int)k;
        Stat s; s.name = name; stats.push_back(s);
        return (int)stats.size() - 1;
    }
    void begin(const char *name, hipStream_t st) override {
        cur = index(name);
        cur_a = get();
        (void)hipEventRecord(cur_a, st);
    }
    void end(hipStream_t st) override {
        hipEvent_t b = get();
        (void)hipEventRecord(b, st);
        pending.push_back({cur, cur_a, b});
    }
    void collect() {
        for (auto &p : pending) {
            float ms = 0.f;
            (void)hipEventSynchronize(p.b);
            (void)hipEventElapsedTime(&ms, p.a, p.b);
            stats[p.idx].launches += 1;
            stats[p.idx].ms += ms;
            pool.push_back(p.a);
            pool.push_back(p.b);
        }
        pending.clear();
    }
};

struct Model {
    game_dims_t d;
    game_config_t cfg;
    int precision = GAME_FP64;
    DevGrid G;
    DevState s[3];
    DevWork W;      /* FP64 form */
    DevWorkF WF;    /* mixed form */
    AdjTables A;    /* reverse tables of the operator adjoints, built on first use */
    PvReverse PR;   /* reverse table of the rhombus PV of the TRSK sums, built with A */
    int adj_ready = 0;
    StepConsts K;
    double idv_dt = -1.0;   /* the dt G.idv holds quotients for */
    hipStream_t stream = nullptr;
    /* a second stream (default priority) for the branches of a step that do
       not depend on each other, forked from and joined back into `stream`
       with the two events; concurrent = 0 runs everything on `stream` */
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    int concurrent = 1;
    std::vector<void *> allocs;
    long long bytes = 0;
    double *partials = nullptr, *h_partials = nullptr;
    double *tmp_s = nullptr;
    EventTimer timer;
    /* HIP graphs of one steady step (not the first), one per assignment of
       the two state buffers to old/new, keyed by the old state's density
       buffer; rebuilt when dt or the configuration changes */
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    const void *graph_key[2] = {nullptr, nullptr};
    double graph_dt = -1.0;
    int use_graphs = 1;
    /* distinct for every model the library has built: buffers made for one
       model are never taken for another at a reused address */
    long long generation = 0;
};

/* copies of one step's inputs and substep results for the reverse sweep
   (FP64 form): the state the step starts from, the state after the first
   substep, and the tendencies of both substeps (rho, rhotheta_v, wind) */
struct StepTape {
    DevState s, s1;
    double *td0_rho, *td0_rt, *td0_wind, *td1_rho, *td1_rt, *td1_wind;
    /* the second substep's mass flux and potential vorticity (FP64 form) */
    double *flux, *pv;
};

extern Model *M;
int dalloc(void **p, size_t n);
/* frees an allocation made by dalloc before the model goes */
void dfree(void *p);
int dzeros(double **dst, size_t count);
int build_adjoint_tables(void);
/* one step of the FP64 form on the resident states, eager launches; with a
   tape, the step's inputs and substep results are copied into it */
int step_fp64(double dt, int first, const StepTape *tape);

}  // namespace rt

#endif
