// Kernel argument blocks and launch declarations of the cfd2d HIP library.
#pragma once
#include "common.h"
#include "state.h"

// Boundary-condition kinds (boundary.py:19-25)
enum { BC_NOSLIP = 0, BC_SLIP = 1, BC_INFLOW = 2, BC_OUTFLOW = 3, BC_PERIODIC = 4 };

#define CFD_MAX_SCALARS 4
#define REC_NV (10 + 5 * CFD_MAX_SCALARS)
#define CFD_MAX_COMP 8
#define PROJ_BS 1024

// Geometry of one grid.  Row pitches are given separately for u-like
// ((nx+1) rows), v-like ((nx+2) rows) and cell ((nx+2) rows) arrays, so the
// same kernels run on the reference's own layout (pitches ny+2, ny+1, ny+2)
// and on the padded layout of a device context (one pitch for all).
struct Geo {
    int nx, ny;
    int pu, pv, pc;
    int periodic;
    double dx, dy;
};

struct PredArgs {
    const double* u;
    const double* v;
    const uint8_t* fl;       // cell flags, bit 0 = fluid (reference `fluid`, ghosts incl.)
    double* hu;
    double* hv;
    const double* hu_old;
    const double* hv_old;
    double* F;
    double* G;
    const double* T;         // Boussinesq temperature or nullptr
    double gx, gy, Tref;
    double nu0;
    const double* nut;       // turbulent viscosity (cells): stress-form diffusion of nu0 + nut;
                             // with rho: the dynamic viscosity of the two-phase mixture
    const double* rho;       // variable density (VOF) or nullptr
    const double* kappa;     // VOF surface tension: curvature (cells), alpha, sigma
    const double* alpha;
    double sigma;
    double fx, fy;
    const DevState* st;
    int scheme_ab2;
    int ext[4];              // outflow extrapolation per side (left, right, bottom, top)
    const uint8_t* open[4];  // open boundary faces per side
};

struct BCArgs {
    double* u;
    double* v;
    const DevState* st;
    int kind[4];
    double vel[4];
    const double* inflow[4];
    const uint8_t* open[4];
};

struct ScalarArgs {
    const double* phi_in;
    double* phi_out;
    double* out;             // standalone transport: right-hand side
    const double* u;
    const double* v;
    const uint8_t* fl;
    const double* diff;      // nullptr: constant D0
    double D0;
    int limiter;
    int obs_t;
    double obs_v;
    int has_solid;
    const int64_t* bt[4];
    const double* bv[4];
    const DevState* st;
};

struct RecordArgs {
    const double* u;
    const double* v;
    const uint8_t* fl;
    const double* p;
    int has_solid;
    int nscalar;
    const double* phi[CFD_MAX_SCALARS];
    const double* diffs[CFD_MAX_SCALARS];
    double D0[CFD_MAX_SCALARS];
    int wallflux[CFD_MAX_SCALARS];
    const int64_t* bt[CFD_MAX_SCALARS][4];
    const double* bv[CFD_MAX_SCALARS][4];
    DevState* st;
    double* part;
    unsigned* flags;         // per-block publish flags (gather_block0)
    SpinGuard spin;          // bound of block 0's wait
    StepRecord* ring;
    int ring_len;
    int scheme_ab2;
    int check_t_end;
    double t_end;
    double steady_tol;
    long long max_steps;
};

#define MG_MAX_LEVELS 16
// One multigrid level: cell arrays of pitch P with a zero ghost ring.
struct MGLevelDev {
    int nx, ny, P;
    int grid;                // 1: spread over all blocks; 0: block 0 alone
    const float* diag;       // level >= 1: Galerkin diagonal (0 = no fluid); level 0: unused
    const float* cE;         // coupling magnitude to the east neighbour (periodic wrap at nx)
    const float* cN;         // coupling magnitude to the north neighbour
    float* b;
    float* x;
    float* t;
};

struct MGDev {
    int nlev;                // levels 0..nlev-1; the last one is solved with cinv
    float omega, wc;
    int ncoarse;             // cells of the coarsest level (nx * ny)
    const float* cinv;       // dense inverse over the coarsest cells (row-major i*ny+j)
    MGLevelDev lev[MG_MAX_LEVELS];
};

// Pressure-projection arguments.  Cell arrays (x, b, r, q, d0, d1, p, code,
// comp) use pitch g.pc and are zero outside the fluid interior.
struct ProjArgs {
    const double* F;
    const double* G;
    double* u;
    double* v;
    double* p;
    const uint8_t* code;     // bits 0-5: index into the diagonal table (0 = not fluid),
                             // bit 6: Dirichlet x-face, bit 7: Dirichlet y-face
    const uint8_t* comp;     // Neumann component 1..ncomp, 0 = none
    const double* lut;       // [0..63] diagonal, [64..127] its reciprocal
    const double* comp_count;
    int ncomp;
    double* x;
    double* b;
    double* r;
    double* q;
    double* d0;
    double* d1;
    float* r32;              // mixed precision work vectors
    float* q32;
    float* d32_0;
    float* d32_1;
    double* rold;            // flexible CG: previous residual
    const MGDev* mg;         // multigrid hierarchy (device memory) for PROJ_MG
    double* sync;            // grid-sync lines: gridDim.x + 2 lines of CFD_SLOT doubles
    SpinGuard spin;          // bound of the grid syncs' waits
    DevState* st;
    double dt_override;      // > 0: use this dt (initial projection with dt = 1)
    const double* rhs_in;    // standalone solve: right-hand side (cell layout) instead of div(F,G)/dt
    int correct;             // 0: solve only (no pressure write-back correction of u, v)
    double rtol;
    int maxiter;
    int zero_p_after;        // initial projection: p := 0 afterwards
    int end_always;          // learned solver's end launches: act even when the solve failed
    int method;              // PROJ_PCG64, PROJ_MG, PROJ_LEARNED
    double reliable_delta;   // mixed: true-residual replacement when ||r|| drops by this factor
    struct FcgState* fcg;    // PROJ_LEARNED: state of the host-driven flexible CG
};

enum { PROJ_PCG64 = 0, PROJ_MG = 1, PROJ_LEARNED = 2 };

// State of the learned flexible CG, whose iterations are separate kernel
// launches (device memory; the reductions finish in the last block to arrive).
struct FcgState {
    double bnorm, atol, rr, rho, rho_old, beta, alpha;
    double cm;               // constant-mode coefficient of the current coarse solve
    double s_in;             // RMS of the current coarse right-hand side (network input scale)
    double irho, irho_old, ibeta, ialpha;   // teacher: flexible CG on the coarse level
    int k, first, done, pending, iters, info, reps, maxiter;
    int net_off;             // the network is switched off for this context
    int skip_net;            // done || net_off: the network's kernels idle
    double rr0;              // r.r at the start of the solve
    long long harvested, hidx;
    unsigned ticket[4];
};

// The flexible CG's state as the host reads it after a batch of iterations:
// k_lp_fin copies the state into this record in pinned host memory, then
// stores the batch's sequence number; the host waits for that number instead
// of copying the state and synchronising the stream.
struct LpPoll {
    FcgState s;
    unsigned seq;
    unsigned pad;
};

// Turbulent viscosity and diffusivities the time-step control reads
// (solver.py:216-218, 249-251, models.py:204-208); nut == nullptr: laminar.
struct TurbDt {
    const double* nut;
    const double* dk;        // k-omega SST diffusivities of k and omega, or nullptr
    const double* dw;
    const uint8_t* fl;
    const double* rho;       // variable density: nut holds mu, the limits use mu / rho
};

// kernels (stencil.hip)
template <int VIS>
__global__ void k_momentum_tendencies(const double* u, const double* v, const uint8_t* fl, Geo g,
                                      double nu0, const double* nuf, double gamma, double* hu,
                                      double* hv);
__global__ void k_predictor(PredArgs a, Geo g);
__global__ void k_copy_predictor(PredArgs a, Geo g);
// domain boundary conditions: work-group 0 and bc_domain_blocks(nx) - 1 more of
// 256 threads, BC_COLS bottom / top columns each
constexpr int BC_COLS = 64;
// threads per work-group of the one-pass reductions k_dt_gamma and k_record,
// one work-group per CU (1024 measured slower: 18 -> 26 and 29 -> 45 us at
// 1024^2, the block reductions over 16 waves and block 0's gather)
constexpr int RED_THREADS = 256;
inline int bc_domain_blocks(int nx) { return 1 + (nx > 2 ? (nx - 2 + BC_COLS - 1) / BC_COLS : 0); }
__global__ void k_bc_domain(BCArgs a, Geo g);
__global__ void k_obstacle_bc(double* u, double* v, const uint8_t* fl, Geo g, const DevState* st);
__global__ void k_dt_gamma(const double* u, const double* v, Geo g, StepParams prm, TurbDt td,
                           DevState* st, double* part, unsigned* flags, SpinGuard spin);
template <bool CONST_D, int MODE>
__global__ void k_scalar_transport(ScalarArgs a, Geo g);
__global__ void k_fill_ghosts(ScalarArgs a, Geo g);
template <int NS>
__global__ void k_record(RecordArgs a, Geo g);
__global__ void k_zero_u32(unsigned* p, int n);
// pcg.hip
__global__ void k_project(ProjArgs a, Geo g);
