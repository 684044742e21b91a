// fx_newton.h -- the device-resident Newton iteration of fx_static_step
// (fx_newton.hip): the convergence decision and the iteration's report stay
// on the device until one read at the iteration's end.
#pragma once

#include "fx_internal.h"

namespace fx {

// The Newton iteration's decision, written by k_nr_check.
struct NrDev {
    int32_t skip;  // 1: ||r|| / ||f_ext|| < 1e-3 (converged), the rest of the iteration does nothing
    int32_t pad;
    double rel;
};

// What the host reads once per iteration (pinned host memory).
struct NrReport {
    int32_t skip;
    int32_t dev_code;  // element status (fx_status of the first failure, 0 if none)
    int64_t bad_element;
    unsigned long long max_diag_bits;  // max_diag of the iteration's assembly
    double rel;
    PcgScalars pcg;  // the solve's scalars after the queued iterations
};

class NewtonGraph {
  public:
    struct FrontArgs {
        bool graph;            // capture the front once and replay it (false: plain launches)
        bool nlgeom;
        bool solve;            // start the FP64 PCG in the front (false: the host solves)
        const fx_solve_opts *opts;
        const uint8_t *mask;   // BC mask of the masked norms
        int64_t n_unique;      // penalty records grouped per DOF
        const int32_t *ud_dof, *ud_ptr, *rec_order;
        const double *target;  // BC targets at the increment's time
    };
    explicit NewtonGraph(fx_mesh *m);
    ~NewtonGraph();
    NewtonGraph(const NewtonGraph &) = delete;
    NewtonGraph &operator=(const NewtonGraph &) = delete;
    void launch_front(const FrontArgs &a);
    // k_nr_post (with_update: du = x, u += du when the solve stopped
    // converged) and the report
    void launch_post(bool with_update);
    const NrReport &wait();  // one stream synchronisation
    const int32_t *skip_flag() const { return &nr.p->skip; }

  private:
    fx_mesh *m;
    DBuf<NrDev> nr;
    DBuf<double> norms;
    hipGraphExec_t front = nullptr;
    NrReport *report = nullptr;
};

}  // namespace fx
