// fx_newton.hip -- the device-resident Newton iteration of fx_static_step.
//
// Reference: src/solver/step/static_step.rs:99-210, one Newton-Raphson
// iteration of an increment: internal force at u, residual f_ext - f_int,
// its norm over the unconstrained DOFs against the external force's
// (converged when the ratio is below 1e-3), tangent K, penalty BCs, the
// linear solve, u += du.
//
// On the device, per iteration, with one host read at its end:
//   front graph (captured once per step): element f_int + gather, residual,
//     masked norms, k_nr_check (the convergence decision, kept on the
//     device), element K + assembly, penalty with 1e6 x max_diag read on the
//     device, Jacobi, the PCG start (stopped at once when converged);
//   the PCG's captured iteration graphs, as many iterations as the previous
//     solve of the step needed, queued without a status read (every launch
//     past the stop rule returns at once);
//   k_nr_post (du = x, u += du when the solve stopped converged and the
//     iteration had not) and k_nr_report (the decision, the solve's scalars,
//     the element status and max_diag into pinned host memory);
//   one stream synchronisation, then the host reads the report.
// A solve still running after the queued iterations finishes with the
// adaptive chunked loop of pcg_solve, and the host adds du itself.  The
// arithmetic of every kernel is the eager loop's: results are bit for bit
// the same (tests/test_gpu_newton.py).
#include "fx_internal.h"
#include "fx_newton.h"

namespace fx {
namespace {

// rel = ||r|| / ||f_ext|| over the unconstrained DOFs (||r|| when
// ||f_ext|| <= 1e-6), static_step.rs:150-160; converged below 1e-3.
__global__ void k_nr_check(const double *__restrict__ norms2, NrDev *__restrict__ nr) {
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    const double res_norm = sqrt(norms2[0]), f_norm = sqrt(norms2[1]);
    const double rel = f_norm > 1e-6 ? res_norm / f_norm : res_norm;
    nr->rel = rel;
    nr->skip = rel < 1e-3 ? 1 : 0;
}

// du = x and u += 1.0 * du (the eager loop's copy and axpy) when the
// iteration had not converged and its solve stopped converged.
__global__ void k_nr_post(int64_t n, const double *__restrict__ x, double *__restrict__ du, double *__restrict__ u,
                          const PcgScalars *__restrict__ sc, const NrDev *__restrict__ nr) {
    if (nr->skip || sc->status != 1)
        return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double d = x[i];
        du[i] = d;
        u[i] += 1.0 * d;
    }
}

__global__ void k_nr_report(const NrDev *__restrict__ nr, const PcgScalars *__restrict__ sc,
                            const DevStatus *__restrict__ st, const unsigned long long *__restrict__ md_bits,
                            NrReport *__restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    NrReport r;
    r.skip = nr->skip;
    r.rel = nr->rel;
    r.pcg = *sc;
    r.dev_code = st->code;
    r.bad_element = st->bad_element;
    r.max_diag_bits = *md_bits;
    *out = r;
    __threadfence_system();
}

inline unsigned grid1(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

NewtonGraph::NewtonGraph(fx_mesh *mesh) : m(mesh) {
    nr.alloc(1);
    norms.alloc(2);
    FX_HIP(hipHostMalloc((void **)&report, sizeof(NrReport), hipHostMallocDefault));
    *report = NrReport{};
}

NewtonGraph::~NewtonGraph() {
    // an error can leave the step with launches still queued: the front
    // graph and k_nr_report (which writes the pinned report) must finish
    // before either is released
    (void)hipStreamSynchronize(m->ctx->stream);
    if (front)
        (void)hipGraphExecDestroy(front);
    if (report)
        (void)hipHostFree(report);
}

// The front of an iteration on the context stream: everything up to the
// PCG start.  Captured on its first use; the buffers it names stay fixed for
// the step.
void NewtonGraph::launch_front(const FrontArgs &a) {
    hipStream_t s = m->ctx->stream;
    auto issue = [&]() {
        m->k_skip = nullptr;
        // 2. internal force at u (coordinates X + u under NLGEOM); the
        //    thermal force is exactly zero (T never changes)
        launch_reset_status(m);
        launch_element(m, false, true, a.nlgeom);
        launch_gather_f(m);
        // 3./4. residual and the BC-masked norms -> the decision
        launch_residual(m, m->f_ext.p, m->f_int.p, m->rhs.p, m->n_dof);
        launch_masked_norms(m, m->rhs.p, m->f_ext.p, a.mask, norms.p);
        hipLaunchKernelGGL(k_nr_check, dim3(1), dim3(64), 0, s, norms.p, nr.p);
        // 5. tangent stiffness, 6. penalty BCs: skipped on the device when
        //    the iteration has converged (K and max_diag keep the values of
        //    the last assembly, as when the host loop stops before them)
        m->k_skip = &nr.p->skip;
        launch_element(m, true, false, a.nlgeom);
        launch_assemble(m);
        launch_penalty_dev(m, a.n_unique, a.ud_dof, a.ud_ptr, a.target, a.rec_order, m->rhs.p, m->u.p);
        m->k_skip = nullptr;
        // 7. the solve's start (stopped at once when converged)
        if (a.solve)
            pcg_start_dev(m, m->rhs.p, *a.opts, &nr.p->skip);
        FX_HIP(hipGetLastError());
    };
    if (!a.graph) {
        issue();
        return;
    }
    if (!front) {
        hipGraph_t g = nullptr;
        FX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            issue();
        } catch (...) {  // end the capture before the error leaves
            m->k_skip = nullptr;
            if (hipStreamEndCapture(s, &g) == hipSuccess && g)
                (void)hipGraphDestroy(g);
            throw;
        }
        FX_HIP(hipStreamEndCapture(s, &g));
        const hipError_t e = hipGraphInstantiate(&front, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        FX_HIP(e);
    }
    FX_HIP(hipGraphLaunch(front, s));
}

void NewtonGraph::launch_post(bool with_update) {
    hipStream_t s = m->ctx->stream;
    if (with_update)
        hipLaunchKernelGGL(k_nr_post, dim3(grid1(m->n_dof, 256)), dim3(256), 0, s, m->n_dof, m->x.p, m->du.p,
                           m->u.p, m->pcg.p, nr.p);
    hipLaunchKernelGGL(k_nr_report, dim3(1), dim3(64), 0, s, nr.p, m->pcg.p, m->status.p, m->max_diag_bits.p,
                       report);
    FX_HIP(hipGetLastError());
}

const NrReport &NewtonGraph::wait() {
    FX_HIP(hipStreamSynchronize(m->ctx->stream));
    ++t_host_syncs;
    return *report;
}

}  // namespace fx
