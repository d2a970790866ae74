// Device-resident per-step scalars and records shared by all step kernels.
#pragma once
#include <stdint.h>

// Scalars of the running simulation.  Written by the step kernels, read by
// the host only through records (no per-step host synchronisation).
struct DevState {
    double t, dt, dt_old, gamma, ab_a, ab_b;
    double umax, vmax;
    double residual;       // max |u_new - u_old| / dt over fluid faces
    double cg_relres;      // last solve: final ||r|| / ||b|| (recursive or true, see solver)
    double cg_bnorm;
    double flow_force;
    long long n;
    int has_old;           // AB2 history (hu_old, hv_old, dt_old) valid
    int cg_iters;          // last solve: iterations (SciPy callback count)
    int cg_info;           // 0 converged, >0 iterations without convergence
    int cg_outer;          // mixed precision: true-residual replacements
    int done;              // run finished (t_end, max_steps, steady_tol, error): later steps are no-ops
    int error;             // instability detected (solver.py:395-399)
    double nut_max;        // max of the turbulent viscosity this step (0 when laminar)
    double flow_du;        // flow-rate control: this step's velocity shift
    double vof_nsub;       // VOF: sub-steps of this step's alpha transport
    double vof_sum, vof_front, vof_height;   // VOF monitors of the step
};

#define CFD_MAX_MON 36
// One record per completed step (cfd2d/solver.py:427-443 plus solver data).
struct StepRecord {
    double t, dt, gamma, residual, divergence, ke;
    double cg_relres;
    double mon[CFD_MAX_MON];   // model monitors in registration order
    long long n;
    int cg_iters, cg_info, cg_outer, nmon;
};

// Host-filled per-step parameters (constant between steps unless changed).
struct StepParams {
    double nu, tau;
    int scheme_ab2;        // 1 = AB2, 0 = Euler
    int gamma_auto;
    double gamma_fixed;
    double dt_fixed;       // > 0: fixed step
    double dt_model;       // min over model dt limits (inf if none)
    double t_end;          // > 0 with clamp: shorten the last step to hit t_end
    int clamp_t_end;
    int pad0;
    double fx, fy;         // body force (flow_force added on device)
};
