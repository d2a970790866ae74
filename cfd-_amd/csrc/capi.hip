// C ABI of the cfd2d MI355X library (include/cfd2d_hip.h): device context,
// reference-layout uploads/downloads, and the launch sequence of one
// projection step.  One translation unit holds the whole library.
#include "cfd_kernels.h"
#include "../../include/cfd2d_hip.h"

#include <algorithm>
#include <cstddef>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "stencil.hip"
#include "pcg.hip"
#include "unet.hip"
#include "lpcg.hip"
#include "ens.hip"
#include "monitors.hip"
#include "turb.hip"
#include "flow.hip"
#include "vof.hip"

// Staging of cfd_download_derived_noalloc: a fixed device array per GPU
// (each device holds its own copy until the process ends).  Host code never
// passes `scratch` itself to a kernel or a copy (in host code the name is a
// stand-in, not a device address): scratch_on() fetches this device's address
// once with hipGetSymbolAddress, and every use goes through that pointer.
static __device__ __align__(256) double scratch[CFD_DERIVED_NOALLOC_MAX];

namespace {

// last error message of the calling thread (a plain array: nothing to destroy at exit)
thread_local char g_err[512];

int set_err(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

constexpr int SCRATCH_MAX_DEV = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[SCRATCH_MAX_DEV] = {};

#define HIPC(expr)                                                                     \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return set_err(CFD_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

constexpr int RING = 8192;
constexpr int BS2D = 256;

struct ScalarDev {
    cfd_scalar_desc desc;
    double* phi[2];
    int64_t* bt[4];
    double* bv[4];
};

}  // namespace

struct cfd_ctx {
    int device = 0;              // HIP device the context was created on
    unsigned* spin_err = nullptr;        // set by a device wait that exceeded spin_limit
    long long spin_limit = 200000000;    // 2 s in ticks of the 100 MHz real-time counter
    int nx, ny, P;
    double dx, dy;
    int periodic;
    int has_solid;
    int kind[4];
    double vel[4];
    int ncomp;
    Geo geo;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    int num_cu = 0;
    int proj_blocks = 0, proj_bs = 0;
    int red_blocks = 0;
    // device arrays (pitch P, nx+2 rows)
    double *u = nullptr, *v = nullptr, *p = nullptr, *F = nullptr, *G = nullptr;
    double* H[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};   // [u/v][buffer]
    double *x = nullptr, *b = nullptr, *r = nullptr, *q = nullptr, *d0 = nullptr, *d1 = nullptr;
    double *tmp0 = nullptr, *tmp1 = nullptr;
    float *r32 = nullptr, *q32 = nullptr, *e0 = nullptr, *e1 = nullptr;
    uint8_t *fl = nullptr, *code = nullptr, *comp = nullptr;
    double *lut = nullptr, *comp_count = nullptr;
    double* inflow[4] = {nullptr, nullptr, nullptr, nullptr};
    uint8_t* open[4] = {nullptr, nullptr, nullptr, nullptr};
    double *sync = nullptr, *part = nullptr;
    unsigned* flags = nullptr;   // per-block publish flags of the one-pass reductions
    DevState* st = nullptr;
    StepRecord* ring = nullptr;
    std::vector<ScalarDev> scalars;
    // function objects on the device (monitors.hip)
    struct {
        int nsample = 0, every = 1, cap = 0;
        int* src = nullptr;
        long long* base = nullptr;
        double *frac = nullptr, *times = nullptr, *vals = nullptr;
        int* count = nullptr;
        std::vector<void*> mem;
    } probes;
    struct {
        int on = 0;
        double t_start = 0.0;
        double *sum = nullptr, *sum2 = nullptr, *T = nullptr;
    } favg;
    long long* mon_last_n = nullptr;
    double* derived = nullptr;   // (nx+1) x (ny+1) node field for reports
    // flow-rate control (flow.hip): start-of-step velocity, kept pressure
    struct {
        double *u0 = nullptr, *v0 = nullptr, *pkeep = nullptr;
        unsigned* ticket = nullptr;
        unsigned long long* acc = nullptr;
        double nfaces = 0.0;
    } flow;
    // two-phase VOF (vof.hip); cell arrays (nx+2) x P; alpha lives in alpha[0]
    // between steps (the sub-step launches come in an even number)
    struct {
        int on = 0;
        double *alpha[2] = {nullptr, nullptr}, *rho = nullptr, *mu = nullptr;
        double *kappa = nullptr, *cE = nullptr, *cN = nullptr, *diag = nullptr;
        VMG mg;                  // multigrid levels of the variable solve (nlev 0: Jacobi)
        int* pinned = nullptr;
        int npinned = 0;
        double* part = nullptr;              // monitors: 3 partials per block
        unsigned* ticket = nullptr;          // last-block tickets of k_vof_nsub, k_vof_monitors
        unsigned long long* acc = nullptr;   // max |u|, max |v|
        double rho1 = 1, rho2 = 1, mu1 = 0, mu2 = 0, sigma = 0, c_alpha = 0, cfl = 0.25, a_in = 0;
    } vof;
    // turbulence model (turb.hip); cell arrays (nx+2) x P
    struct {
        int model = TURB_NONE;
        double nu = 0.0;
        double *nut = nullptr, *l2 = nullptr, *d = nullptr, *F1 = nullptr, *cross = nullptr;
        double *dk = nullptr, *dw = nullptr, *wwall = nullptr;
        uint8_t* wadj = nullptr;
        int sid_k = -1, sid_w = -1;
    } turb;
    double* rold = nullptr;
    MGDev* mg_dev = nullptr;     // device copy of the hierarchy description
    MGDev mg_host;               // host copy (level sizes and array pointers)
    int mg_ready = 0;
    // learned flexible CG (lpcg.hip)
    int l_ready = 0;
    int l_level = 0, l_mode = 0, l_cycles = 0, l_fallback = 1, l_graph = 1, l_maxiter = 200;
    int l_fused = 0, l_lc = 0;   // coarse levels l_lc .. l_level in one k_lv_coarse launch
    int l_vlds = 0;              // mode V-cycle: levels l_lc .. coarsest in LDS (k_lv_vcoarse)
    int l_pair = 0;              // levels l_lc - 2 and l_lc - 1 in one launch per leg
    int l_legs = 0;              // levels 0 .. l_lc - 1: one launch per level and leg (2: not
                                 // the fine up-leg)
    int l_sweeps = 2;            // smoothing sweeps per leg on those levels (this solve)
    int l_sweeps_cfg = 2;        // 2, 3, 4, or -1: 2 or 3 by the last solves' iterations
    int l_low_run = 0;           // consecutive 3-sweep solves of at most 2 iterations
    VcArgs l_vargs;
    int l_x0 = 0, l_hist = 0;    // extrapolated initial iterate; x_prev holds the last solution
    double* xprev = nullptr;
    int l_H = 0, l_W = 0;
    float l_sigma = 1.0f, l_D = 1.0f;
    double l_1a1 = 0.0;
    cfd_unet* l_net = nullptr;
    TeachVec l_tv = {};
    int l_tv_level = -1;
    float *l_hin = nullptr, *l_hout = nullptr;
    long long l_hcap = 0;
    float* diag0f = nullptr;
    FcgState* fcg = nullptr;
    LpPoll* poll_h = nullptr;       // pinned record of the solve's state (host address)
    LpPoll* poll_d = nullptr;       // the same record as the device addresses it
    unsigned poll_seq = 0;          // sequence number of the last end launches
    double* l_part = nullptr;
    hipGraph_t l_graph_g[3] = {};    // one captured iteration per sweep count 2, 3, 4
    hipGraphExec_t l_graph_x[3] = {};
    int l_prev_iters = 0;
    long long l_fallbacks = 0, l_solves = 0, l_iters_total = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // optional timing of every projection launch (HIP events on the context stream)
    int prof_on = 0;
    size_t prof_used = 0;
    std::vector<hipEvent_t> prof_ev;
    std::vector<void*> allocs;
};

namespace {

// zeroed device array; the fill runs on the context stream and is complete on
// return: a plain hipMemset goes to the null stream, which the context's
// non-blocking stream does not wait for, so a kernel issued right after a
// lazy allocation could see its output cleared behind it
template <typename T>
int dalloc(cfd_ctx* c, T** p, size_t n) {
    void* q = nullptr;
    const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    HIPC(hipMalloc(&q, bytes));
    if (c->stream) {
        HIPC(hipMemsetAsync(q, 0, bytes, c->stream));
        HIPC(hipStreamSynchronize(c->stream));
    } else {
        HIPC(hipMemset(q, 0, bytes));
        HIPC(hipDeviceSynchronize());
    }
    c->allocs.push_back(q);
    *p = (T*)q;
    return 0;
}

#define TRY(expr)                 \
    do {                          \
        int rc_ = (expr);         \
        if (rc_ != 0) return rc_; \
    } while (0)

size_t cells(const cfd_ctx* c) { return (size_t)(c->nx + 2) * c->P; }

dim3 grid2d(const cfd_ctx* c) { return dim3((c->ny + 2 + BS2D - 1) / BS2D, c->nx + 2); }

int shape_of(const cfd_ctx* c, int field, int* rows, int* cols, int* off_i, int* off_j) {
    *off_i = 0;
    *off_j = 0;
    switch (field) {
    case CFD_FIELD_U: case CFD_FIELD_F: case CFD_FIELD_HU: case CFD_FIELD_HU_OLD:
        *rows = c->nx + 1; *cols = c->ny + 2; return 0;
    case CFD_FIELD_V: case CFD_FIELD_G: case CFD_FIELD_HV: case CFD_FIELD_HV_OLD:
        *rows = c->nx + 2; *cols = c->ny + 1; return 0;
    case CFD_FIELD_P:
        *rows = c->nx + 2; *cols = c->ny + 2; return 0;
    case CFD_FIELD_X: case CFD_FIELD_XPREV:
        *rows = c->nx; *cols = c->ny; *off_i = 1; *off_j = 1; return 0;
    case CFD_FIELD_NUT: case CFD_FIELD_DIFF_K: case CFD_FIELD_DIFF_OMEGA:
    case CFD_FIELD_ALPHA: case CFD_FIELD_RHO: case CFD_FIELD_MU:
        *rows = c->nx + 2; *cols = c->ny + 2; return 0;
    default:
        return set_err(CFD_ERR_ARG, "unknown field id %d", field);
    }
}

// the address of `scratch` on device `dev` (fetched once per device, with that
// device current; the caller's current device is restored)
int scratch_on(int dev, double** out) {
    if (dev < 0 || dev >= SCRATCH_MAX_DEV)
        return set_err(CFD_ERR_LIMIT, "device %d: the allocation-free staging serves devices "
                       "0..%d", dev, SCRATCH_MAX_DEV - 1);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        int cur = 0;
        HIPC(hipGetDevice(&cur));
        if (cur != dev) HIPC(hipSetDevice(dev));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) HIPC(hipSetDevice(cur));
        if (e != hipSuccess || !p)
            return set_err(CFD_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                           hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

SpinGuard spin_guard(const cfd_ctx* c) { return SpinGuard{c->spin_err, c->spin_limit}; }

// a device wait that gave up leaves the state invalid; e is the error word
// (read here after synchronising the stream, unless the caller passes it)
int check_spin(cfd_ctx* c, const unsigned* word = nullptr) {
    unsigned e = 0;
    if (word) e = *word;
    else HIPC(hipMemcpy(&e, c->spin_err, sizeof(e), hipMemcpyDeviceToHost));
    if (e)
        return set_err(CFD_ERR_STATE, "a grid synchronisation wait on the device exceeded its "
                       "limit of %.3g s and gave up: the device state is invalid "
                       "(cfd_set_spin_limit clears the error)", c->spin_limit * 1e-8);
    return 0;
}

int read_state(cfd_ctx* c, DevState* h) {
    unsigned e = 0;
    HIPC(hipMemcpyAsync(h, c->st, sizeof(DevState), hipMemcpyDeviceToHost, c->stream));
    HIPC(hipMemcpyAsync(&e, c->spin_err, sizeof(e), hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    TRY(check_spin(c, &e));
    return 0;
}

double* field_ptr(cfd_ctx* c, int field, int par) {
    switch (field) {
    case CFD_FIELD_U: return c->u;
    case CFD_FIELD_V: return c->v;
    case CFD_FIELD_P: return c->p;
    case CFD_FIELD_F: return c->F;
    case CFD_FIELD_G: return c->G;
    // parity par: H[.][par] is written by the next step, H[.][par^1] holds the last step's
    case CFD_FIELD_HU: return c->H[0][par ^ 1];
    case CFD_FIELD_HV: return c->H[1][par ^ 1];
    case CFD_FIELD_HU_OLD: return c->H[0][par ^ 1];
    case CFD_FIELD_HV_OLD: return c->H[1][par ^ 1];
    case CFD_FIELD_X: return c->x;
    case CFD_FIELD_XPREV: return c->xprev;
    case CFD_FIELD_NUT: return c->turb.nut;
    case CFD_FIELD_DIFF_K: return c->turb.dk;
    case CFD_FIELD_DIFF_OMEGA: return c->turb.dw;
    case CFD_FIELD_ALPHA: return c->vof.alpha[0];
    case CFD_FIELD_RHO: return c->vof.rho;
    case CFD_FIELD_MU: return c->vof.mu;
    default: return nullptr;
    }
}

// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------
BCArgs bc_args(cfd_ctx* c, bool check_done) {
    BCArgs a;
    a.u = c->u;
    a.v = c->v;
    a.st = check_done ? c->st : nullptr;
    for (int s = 0; s < 4; ++s) {
        a.kind[s] = c->kind[s];
        a.vel[s] = c->vel[s];
        a.inflow[s] = c->inflow[s];
        a.open[s] = c->open[s];
    }
    return a;
}

int launch_bc(cfd_ctx* c, bool check_done) {
    hipLaunchKernelGGL(k_bc_domain, dim3(bc_domain_blocks(c->nx)), dim3(256), 0, c->stream,
                       bc_args(c, check_done), c->geo);
    HIPC(hipGetLastError());
    if (c->has_solid || c->periodic) {
        hipLaunchKernelGGL(k_obstacle_bc, grid2d(c), dim3(BS2D), 0, c->stream, c->u, c->v, c->fl,
                           c->geo, check_done ? c->st : (const DevState*)nullptr);
        HIPC(hipGetLastError());
    }
    return 0;
}

PredArgs pred_args(cfd_ctx* c, const cfd_step_params* prm, int par) {
    PredArgs a;
    std::memset(&a, 0, sizeof(a));
    a.u = c->u;
    a.v = c->v;
    a.fl = c->fl;
    a.hu = c->H[0][par];
    a.hv = c->H[1][par];
    a.hu_old = c->H[0][par ^ 1];
    a.hv_old = c->H[1][par ^ 1];
    a.F = c->F;
    a.G = c->G;
    a.T = nullptr;
    for (auto& s : c->scalars) {
        if (s.desc.boussinesq) {
            a.T = s.phi[par];
            a.gx = s.desc.gx;
            a.gy = s.desc.gy;
            a.Tref = s.desc.t_ref;
        }
    }
    a.nu0 = prm ? prm->nu : 0.0;
    a.nut = c->turb.model != TURB_NONE ? c->turb.nut : nullptr;
    if (c->vof.on) {
        a.nut = c->vof.mu;
        a.rho = c->vof.rho;
        if (c->vof.sigma > 0.0) {
            a.kappa = c->vof.kappa;
            a.alpha = c->vof.alpha[0];
            a.sigma = c->vof.sigma;
        }
    }
    a.fx = prm ? prm->fx : 0.0;
    a.fy = prm ? prm->fy : 0.0;
    a.st = c->st;
    a.scheme_ab2 = prm ? prm->scheme_ab2 : 0;
    for (int s = 0; s < 4; ++s) {
        a.ext[s] = c->kind[s] == CFD_BC_OUTFLOW;
        a.open[s] = c->open[s];
    }
    return a;
}

ProjArgs proj_args(cfd_ctx* c, int method, double rtol, int maxiter) {
    ProjArgs a;
    std::memset(&a, 0, sizeof(a));
    a.F = c->F;
    a.G = c->G;
    a.u = c->u;
    a.v = c->v;
    a.p = c->p;
    a.code = c->code;
    a.comp = c->comp;
    a.lut = c->lut;
    a.comp_count = c->comp_count;
    a.ncomp = c->ncomp;
    a.x = c->x;
    a.b = c->b;
    a.r = c->r;
    a.q = c->q;
    a.d0 = c->d0;
    a.d1 = c->d1;
    a.r32 = c->r32;
    a.q32 = c->q32;
    a.d32_0 = c->e0;
    a.d32_1 = c->e1;
    a.sync = c->sync;
    a.spin = spin_guard(c);
    a.rold = c->rold;
    a.mg = c->mg_dev;
    a.st = c->st;
    a.rtol = rtol;
    a.maxiter = maxiter;
    a.method = method;
    a.correct = 1;
    a.reliable_delta = 0.1;
    return a;
}

int run_learned(cfd_ctx* c, ProjArgs a);

// one cooperative launch of k_project (timed when profiling is on and prof)
int launch_coop(cfd_ctx* c, ProjArgs a, bool prof = true) {
    const int words = (c->proj_blocks + 2) * CFD_SLOT * 2;
    hipLaunchKernelGGL(k_zero_u32, dim3((words + 255) / 256), dim3(256), 0, c->stream,
                       (unsigned*)c->sync, words);
    HIPC(hipGetLastError());
    Geo g = c->geo;
    void* args[] = {(void*)&a, (void*)&g};
    prof = prof && c->prof_on;
    if (prof) {
        if (c->prof_used + 2 > c->prof_ev.size()) {
            for (int k = 0; k < 256; ++k) {
                hipEvent_t e;
                HIPC(hipEventCreate(&e));
                c->prof_ev.push_back(e);
            }
        }
        HIPC(hipEventRecord(c->prof_ev[c->prof_used], c->stream));
    }
    HIPC(hipLaunchCooperativeKernel((const void*)k_project, dim3(c->proj_blocks),
                                    dim3(c->proj_bs), args, 0, c->stream));
    if (prof) {
        HIPC(hipEventRecord(c->prof_ev[c->prof_used + 1], c->stream));
        c->prof_used += 2;
    }
    return 0;
}

constexpr int VOF_NSUB_MAX = 8;   // even: alpha returns to alpha[0]

VofArgs vof_args(cfd_ctx* c, const double* in, double* out, int sub) {
    VofArgs a;
    std::memset(&a, 0, sizeof(a));
    a.a = in;
    a.out = out;
    a.u = c->u;
    a.v = c->v;
    a.fl = c->fl;
    a.rho = c->vof.rho;
    a.mu = c->vof.mu;
    a.kappa = c->vof.kappa;
    a.st = c->st;
    a.rho1 = c->vof.rho1;
    a.rho2 = c->vof.rho2;
    a.mu1 = c->vof.mu1;
    a.mu2 = c->vof.mu2;
    a.c_alpha = c->vof.c_alpha;
    a.cfl = c->vof.cfl;
    a.a_in = c->vof.a_in;
    a.sub = sub;
    a.nsub_max = VOF_NSUB_MAX;
    a.part = c->vof.part;
    a.ticket = c->vof.ticket;
    a.acc = c->vof.acc;
    return a;
}

// the variable-coefficient pressure solve and correction, one cooperative launch
int launch_coop_var(cfd_ctx* c, ProjArgs a) {
    const int words = (c->proj_blocks + 2) * CFD_SLOT * 2;
    hipLaunchKernelGGL(k_zero_u32, dim3((words + 255) / 256), dim3(256), 0, c->stream,
                       (unsigned*)c->sync, words);
    HIPC(hipGetLastError());
    VarProjArgs va;
    va.p = a;
    va.p.rtol = std::min(a.rtol, 1e-12);   // the reference solves this system directly
    va.rho = c->vof.rho;
    va.cE = c->vof.cE;
    va.cN = c->vof.cN;
    va.diag = c->vof.diag;
    va.pinned = c->vof.pinned;
    va.npinned = c->vof.npinned;
    va.mg = c->vof.mg;
    va.mg.lev[0].cE = c->vof.cE;
    va.mg.lev[0].cN = c->vof.cN;
    va.mg.lev[0].diag = c->vof.diag;
    va.mg.lev[0].b = a.r;
    Geo g = c->geo;
    void* args[] = {(void*)&va, (void*)&g};
    // the multigrid solve runs the levels of up to 4 cells per thread in block
    // 0 behind block barriers: full 1024-thread blocks keep more levels there
    const int bs = va.mg.nlev >= 2 ? PROJ_BS : c->proj_bs;
    HIPC(hipLaunchCooperativeKernel((const void*)k_project_var, dim3(c->proj_blocks), dim3(bs),
                                    args, 0, c->stream));
    return 0;
}

// initial iterate of a time step's constant-density solve: x <- x + theta
// (x - x_prev) when the extrapolated start is set (not for the initial
// projection, which passes dt_override); x_prev keeps the last solution
int x0_extrapolate(cfd_ctx* c, const ProjArgs& a) {
    if (!c->l_x0 || a.dt_override != 0.0) return 0;
    const long n = (long)(c->nx + 2) * c->P;
    hipLaunchKernelGGL(k_x0_extrap, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                       c->x, c->xprev, n, (const DevState*)c->st, c->l_hist);
    HIPC(hipGetLastError());
    c->l_hist = 1;
    return 0;
}

int launch_project(cfd_ctx* c, ProjArgs a) {
    if (c->vof.on) return launch_coop_var(c, a);
    if (a.method != PROJ_PCG64 && a.method != PROJ_MG && a.method != PROJ_LEARNED)
        return set_err(CFD_ERR_ARG, "pressure solver %d is not available", a.method);
    if (a.method == PROJ_MG && !c->mg_ready)
        return set_err(CFD_ERR_STATE, "multigrid solver selected before cfd_set_multigrid");
    if (a.method == PROJ_LEARNED) {
        if (!c->l_ready)
            return set_err(CFD_ERR_STATE, "learned solver selected before cfd_set_learned");
        return run_learned(c, a);
    }
    TRY(x0_extrapolate(c, a));
    return launch_coop(c, a);
}

ScalarArgs scalar_args(cfd_ctx* c, ScalarDev& s, int par) {
    ScalarArgs a;
    std::memset(&a, 0, sizeof(a));
    a.phi_in = s.phi[par];
    a.phi_out = s.phi[par ^ 1];
    a.u = c->u;
    a.v = c->v;
    a.fl = c->fl;
    a.diff = nullptr;
    a.D0 = s.desc.diffusivity;
    a.limiter = s.desc.limiter;
    a.obs_t = s.desc.obstacle_fixed ? 1 : 0;
    a.obs_v = s.desc.obstacle_fixed ? s.desc.obstacle_value : 0.0;
    a.has_solid = c->has_solid;
    for (int k = 0; k < 4; ++k) {
        a.bt[k] = s.bt[k];
        a.bv[k] = s.bv[k];
    }
    a.st = c->st;
    return a;
}

TurbArgs turb_args(cfd_ctx* c, int par, double nu, bool check_done) {
    TurbArgs a;
    std::memset(&a, 0, sizeof(a));
    a.model = c->turb.model;
    a.nu = nu;
    a.u = c->u;
    a.v = c->v;
    a.fl = c->fl;
    a.nut = c->turb.nut;
    a.l2 = c->turb.l2;
    if (c->turb.model == TURB_SST) {
        a.kphi = c->scalars[c->turb.sid_k].phi[par];
        a.wphi = c->scalars[c->turb.sid_w].phi[par];
    }
    a.d = c->turb.d;
    a.F1 = c->turb.F1;
    a.cross = c->turb.cross;
    a.dk = c->turb.dk;
    a.dw = c->turb.dw;
    for (int k = 0; k < 4; ++k) a.kind[k] = c->kind[k];
    a.st = check_done ? c->st : nullptr;
    return a;
}

// nu_t (and the SST diffusivities) from the velocity and k, omega of parity
// par: inner cells, then the ghost ring
int launch_turb_pre(cfd_ctx* c, int par, double nu, bool check_done) {
    const TurbArgs a = turb_args(c, par, nu, check_done);
    hipLaunchKernelGGL(k_turb_cells, dim3((c->ny + 255) / 256, c->nx), dim3(256), 0, c->stream, a,
                       c->geo);
    HIPC(hipGetLastError());
    const int ng = 2 * (c->nx + 2) + 2 * c->ny;
    hipLaunchKernelGGL(k_turb_ghosts, dim3((ng + 255) / 256), dim3(256), 0, c->stream, a, c->geo);
    HIPC(hipGetLastError());
    return 0;
}

// k and omega after the step (parity par -> par ^ 1), ghost fills included
int launch_sst_post(cfd_ctx* c, int par) {
    ScalarDev& K = c->scalars[c->turb.sid_k];
    ScalarDev& W = c->scalars[c->turb.sid_w];
    SstArgs s;
    s.k = scalar_args(c, K, par);
    s.w = scalar_args(c, W, par);
    s.k.diff = c->turb.dk;
    s.w.diff = c->turb.dw;
    s.nut = c->turb.nut;
    s.F1 = c->turb.F1;
    s.cross = c->turb.cross;
    hipLaunchKernelGGL(k_sst_update, grid2d(c), dim3(BS2D), 0, c->stream, s, c->geo);
    HIPC(hipGetLastError());
    hipLaunchKernelGGL(k_fill_ghosts, grid2d(c), dim3(BS2D), 0, c->stream, s.k, c->geo);
    HIPC(hipGetLastError());
    hipLaunchKernelGGL(k_fill_ghosts, grid2d(c), dim3(BS2D), 0, c->stream, s.w, c->geo);
    HIPC(hipGetLastError());
    // the old omega buffer keeps the first fill; the wall cells get omega_wall
    hipLaunchKernelGGL(k_sst_wall, grid2d(c), dim3(BS2D), 0, c->stream, s.w.phi_out,
                       (double*)s.w.phi_in, (const uint8_t*)c->turb.wadj,
                       (const double*)c->turb.wwall, c->geo, (const DevState*)c->st);
    HIPC(hipGetLastError());
    hipLaunchKernelGGL(k_fill_ghosts, grid2d(c), dim3(BS2D), 0, c->stream, s.w, c->geo);
    HIPC(hipGetLastError());
    return 0;
}

// flow-rate control after the correction and the boundary conditions
// (solver.py:372-373, 410-425); u0/v0 hold the velocity of the step's start
int launch_flow_control(cfd_ctx* c, const cfd_step_params* prm) {
    FlowArgs a;
    a.u = c->u;
    a.v = c->v;
    a.fl = c->fl;
    a.st = c->st;
    a.flow_rate = prm->flow_rate;
    a.nfaces = c->flow.nfaces;
    a.part = c->part;
    a.ticket = c->flow.ticket;
    a.acc = c->flow.acc;
    a.u0 = c->flow.u0;
    a.v0 = c->flow.v0;
    hipLaunchKernelGGL(k_flow_bulk, dim3(c->red_blocks), dim3(256), 0, c->stream, a, c->geo);
    HIPC(hipGetLastError());
    hipLaunchKernelGGL(k_flow_shift, grid2d(c), dim3(BS2D), 0, c->stream, a, c->geo);
    HIPC(hipGetLastError());
    if (c->has_solid) {
        // the shifted field is not divergence-free around obstacles: project
        // it again with dt = 1 and keep the step's pressure
        PredArgs pa = pred_args(c, prm, 0);
        hipLaunchKernelGGL(k_copy_predictor, grid2d(c), dim3(BS2D), 0, c->stream, pa, c->geo);
        HIPC(hipGetLastError());
        HIPC(hipMemcpyAsync(c->flow.pkeep, c->p, cells(c) * sizeof(double),
                            hipMemcpyDeviceToDevice, c->stream));
        ProjArgs pj = proj_args(c, prm->solver, prm->rtol, prm->maxiter);
        pj.dt_override = 1.0;
        TRY(launch_project(c, pj));
        HIPC(hipMemcpyAsync(c->p, c->flow.pkeep, cells(c) * sizeof(double),
                            hipMemcpyDeviceToDevice, c->stream));
        c->l_hist = 0;   // the solves no longer form one time series
    }
    TRY(launch_bc(c, true));
    hipLaunchKernelGGL(k_flow_residual, dim3(c->red_blocks), dim3(256), 0, c->stream, a, c->geo);
    HIPC(hipGetLastError());
    return 0;
}


// fill the ghost and solid cells of alpha[k] and set rho, mu (sub -1: always)
int launch_vof_fill(cfd_ctx* c, int k, int sub) {
    VofArgs a = vof_args(c, c->vof.alpha[k], c->vof.alpha[k], sub);
    const int ng = 2 * (c->nx + 2) + 2 * c->ny;
    hipLaunchKernelGGL(k_vof_ghosts, dim3((ng + 255) / 256), dim3(256), 0, c->stream, a, c->geo);
    HIPC(hipGetLastError());
    hipLaunchKernelGGL(k_vof_solids, grid2d(c), dim3(BS2D), 0, c->stream, a, c->geo);
    HIPC(hipGetLastError());
    return 0;
}

// alpha transport after the step (vof.py:243-254) and its monitors
int launch_vof_post(cfd_ctx* c) {
    VofArgs a0 = vof_args(c, c->vof.alpha[0], c->vof.alpha[1], 0);
    hipLaunchKernelGGL(k_vof_nsub, dim3(c->red_blocks), dim3(256), 0, c->stream, a0, c->geo);
    HIPC(hipGetLastError());
    for (int s = 0; s < VOF_NSUB_MAX; ++s) {
        VofArgs a = vof_args(c, c->vof.alpha[s & 1], c->vof.alpha[(s & 1) ^ 1], s);
        hipLaunchKernelGGL(k_vof_advect, grid2d(c), dim3(BS2D), 0, c->stream, a, c->geo);
        HIPC(hipGetLastError());
        TRY(launch_vof_fill(c, (s & 1) ^ 1, s));
    }
    // the properties update fills the ghosts once more (vof.py:205-206): a
    // ghost next to a solid edge cell then takes that cell's new value
    TRY(launch_vof_fill(c, 0, -1));
    VofArgs m = vof_args(c, c->vof.alpha[0], c->vof.alpha[0], -1);
    hipLaunchKernelGGL(k_vof_monitors, dim3(c->red_blocks), dim3(256), 0, c->stream, m, c->geo);
    HIPC(hipGetLastError());
    return 0;
}

int host_parity(cfd_ctx* c, int* par) {
    DevState h;
    TRY(read_state(c, &h));
    *par = (int)(h.n & 1);
    return 0;
}

}  // namespace

#include "unet_net.hip"

namespace {

LvArgs lv_args(cfd_ctx* c) {
    LvArgs a;
    a.mg = c->mg_dev;
    a.code = c->code;
    a.diag0 = c->diag0f;
    a.periodic = c->periodic;
    a.ax = (float)(1.0 / (c->dx * c->dx));
    a.ay = (float)(1.0 / (c->dy * c->dy));
    a.fs = c->fcg;
    a.part = c->l_part;
    a.store_x0 = (c->l_legs && (c->l_vlds || c->l_fused)) ? 0 : 1;
    return a;
}

LegArgs leg_args(cfd_ctx* c, int l) {
    const MGDev& mg = c->mg_host;
    LegArgs g;
    g.L = mg.lev[l];
    g.C = mg.lev[l + 1];
    // the down-leg's t in x, except on level 0 when its up-leg runs as
    // k_lv_post + k_lv_smooth_dots (which read t)
    g.tdown = (c->l_legs == 2 && l == 0) ? g.L.t : g.L.x;
    g.tup = g.tdown;
    g.code = c->code;
    g.diag0 = c->diag0f;
    g.done = &c->fcg->done;
    g.first = &c->fcg->first;
    g.r = c->r;
    g.rold = c->rold;
    g.part = c->l_part;
    g.w = mg.omega;
    g.wc = mg.wc;
    g.ax = (float)(1.0 / (c->dx * c->dx));
    g.ay = (float)(1.0 / (c->dy * c->dy));
    g.periodic = c->periodic;
    return g;
}

long leg_tiles_of(const MGLevelDev& L, int ti, int tj) {
    return (long)((L.nx + ti - 1) / ti) * ((L.ny + tj - 1) / tj);
}

// 16 x 64 tiles when they give at least 512 work-groups
bool leg_big(const MGLevelDev& L) { return leg_tiles_of(L, 16, 64) >= 512; }

unsigned leg_grid(const MGLevelDev& L, bool big, long cap) {
    return (unsigned)std::min(leg_tiles_of(L, big ? 16 : 8, big ? 64 : 32), cap);
}

using LegKernel = void (*)(LegArgs);

template <int NU>
LegKernel leg_down_nu(bool lev0, bool big) {
    return lev0 ? (big ? k_lv_down<true, NU, 16, 64> : k_lv_down<true, NU, 8, 32>)
                : (big ? k_lv_down<false, NU, 16, 64> : k_lv_down<false, NU, 8, 32>);
}

// the fine level's up-leg with the dot partials
template <int NU>
LegKernel leg_up_nu(bool lev0, bool big) {
    return lev0 ? (big ? k_lv_up<true, true, NU, 16, 64> : k_lv_up<true, true, NU, 8, 32>)
                : (big ? k_lv_up<false, false, NU, 16, 64> : k_lv_up<false, false, NU, 8, 32>);
}

// sweeps per leg: 2, 3 or 4 (checked by cfd_set_learned)
LegKernel leg_down_kernel(int nu, bool lev0, bool big) {
    return nu == 4 ? leg_down_nu<4>(lev0, big) : nu == 3 ? leg_down_nu<3>(lev0, big) : leg_down_nu<2>(lev0, big);
}

LegKernel leg_up_kernel(int nu, bool lev0, bool big) {
    return nu == 4 ? leg_up_nu<4>(lev0, big) : nu == 3 ? leg_up_nu<3>(lev0, big) : leg_up_nu<2>(lev0, big);
}

// two coarse levels per leg launch (8 x 32 tiles of the coarser one)
using LegKernel2 = void (*)(LegArgs, LegArgs);
constexpr int LEG2_DOWN_TI = 8, LEG2_DOWN_TJ = 16;   // the down-leg's M tiles
LegKernel2 leg_down2_kernel(int nu) {
    return nu == 3 ? k_lv_down2<3, LEG2_DOWN_TI, LEG2_DOWN_TJ> : k_lv_down2<2, LEG2_DOWN_TI, LEG2_DOWN_TJ>;
}
constexpr int LEG2_UP_TI = 8, LEG2_UP_TJ = 16;   // the up-leg's M tiles
LegKernel2 leg_up2_kernel(int nu) {
    return nu == 3 ? k_lv_up2<3, LEG2_UP_TI, LEG2_UP_TJ> : k_lv_up2<2, LEG2_UP_TI, LEG2_UP_TJ>;
}
unsigned leg_grid2(const MGLevelDev& M, bool up) {
    return (unsigned)std::min(leg_tiles_of(M, up ? LEG2_UP_TI : LEG2_DOWN_TI, up ? LEG2_UP_TJ : LEG2_DOWN_TJ),
                              2048L);
}

FcgVec fcg_vec(cfd_ctx* c) {
    FcgVec v;
    v.x = c->x;
    v.b = c->b;
    v.r = c->r;
    v.rold = c->rold;
    v.q = c->q;
    v.d0 = c->d0;
    v.d1 = c->d1;
    v.lut = c->lut;
    v.nx = c->nx;
    v.pc = c->P;
    v.ax = 1.0 / (c->dx * c->dx);
    v.ay = 1.0 / (c->dy * c->dy);
    return v;
}

#define LV_LAUNCH(kern, blocks, ...)                                                       \
    do {                                                                                   \
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, c->stream, __VA_ARGS__);       \
        HIPC(hipGetLastError());                                                           \
    } while (0)
#define LV_LAUNCH512(kern, blocks, ...)                                                    \
    do {                                                                                   \
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(512), 0, c->stream, __VA_ARGS__);       \
        HIPC(hipGetLastError());                                                           \
    } while (0)

// the kernels of one flexible-CG iteration, in stream order.  Reducing
// kernels run in bands of rows (at most 512 work-groups, whose partial sums the
// next kernel adds); the other sweeps use one cell per thread.
int learned_iteration(cfd_ctx* c) {
    const LvArgs a = lv_args(c);
    const FcgVec v = fcg_vec(c);
    const MGDev& mg = c->mg_host;
    const int K = c->l_level, NL = mg.nlev;
    auto bands = [&](int l) { return std::min(mg.lev[l].nx, 512); };
    auto cells = [&](int l) {
        return (unsigned)(mg.lev[l].nx * ((mg.lev[l].ny + 255) / 256));
    };
    const int nb0 = bands(0), nbK = bands(K);
    const unsigned nimg = (unsigned)(((long)c->l_H * c->l_W + 255) / 256);
    // the legs of the V-cycle on level l, and the level-0 up-leg ending in
    // z with the partials of r.z and r_old.z
    auto down = [&](int l) -> int {
        LV_LAUNCH(k_lv_smooth, cells(l), a, l, 0);
        LV_LAUNCH(k_lv_restrict, cells(l + 1), a, l);
        return 0;
    };
    auto up = [&](int l) -> int {
        LV_LAUNCH(k_lv_post, cells(l), a, l);
        LV_LAUNCH(k_lv_smooth, cells(l), a, l, 0);
        return 0;
    };
    auto up0 = [&]() -> int {
        LV_LAUNCH(k_lv_post, cells(0), a, 0);
        LV_LAUNCH512(k_lv_smooth_dots, nb0, a, v);
        return 0;
    };
    // the flexible-CG steps after z: direction (adding nb_dots partials of r.z
    // and r_old.z), update, recursive-residual test and true residual, end of
    // the iteration
    auto tail = [&](int nb_dots, int net_used) -> int {
        LV_LAUNCH512(k_fcg_dir, nb0, a, v, nb_dots);
        LV_LAUNCH512(k_fcg_update, nb0, a, v, nb0);
        LV_LAUNCH512(k_fcg_true, nb0, a, v, nb0);
        LV_LAUNCH(k_fcg_true_fin, 1, a, nb0, net_used);
        return 0;
    };
    // the legs of levels 0 .. Lc - 1 as one launch per level and leg (tiles
    // of 16 x 64 cells where they give at least 512 work-groups, else 8 x 32);
    // returns the work-group count of the level-0 up-leg (its dot partials)
    const int nu = c->l_sweeps;
    const bool pair = c->l_pair && c->l_vlds && nu <= 3;   // (4 sweeps: the LDS region spills)
    auto legs_down = [&](int Lc) -> int {
        for (int l = 0; l < (pair ? Lc - 2 : Lc); ++l) {
            const LegArgs g = leg_args(c, l);
            const bool big = leg_big(g.L);
            LV_LAUNCH(leg_down_kernel(nu, l == 0, big), leg_grid(g.L, big, 2048), g);
        }
        if (pair) {
            const LegArgs f = leg_args(c, Lc - 2), m = leg_args(c, Lc - 1);
            hipLaunchKernelGGL(leg_down2_kernel(nu), dim3(leg_grid2(m.L, false)), dim3(LEG2_DOWN_BS),
                               0, c->stream, f, m);
            HIPC(hipGetLastError());
        }
        return 0;
    };
    auto legs_up = [&](int Lc, int* nb_dots) -> int {
        if (c->l_legs == 2) {   // the fine level's up-leg as two launches
            for (int l = Lc - 1; l >= 1; --l) {
                const LegArgs g = leg_args(c, l);
                const bool big = leg_big(g.L);
                LV_LAUNCH(leg_up_kernel(2, false, big), leg_grid(g.L, big, 2048), g);
            }
            TRY(up0());
            *nb_dots = nb0;
            return 0;
        }
        if (pair) {
            const LegArgs m = leg_args(c, Lc - 1), f = leg_args(c, Lc - 2);
            hipLaunchKernelGGL(leg_up2_kernel(nu), dim3(leg_grid2(m.L, true)), dim3(LEG2_UP_BS),
                               0, c->stream, m, f);
            HIPC(hipGetLastError());
        }
        for (int l = pair ? Lc - 3 : Lc - 1; l >= 0; --l) {
            const LegArgs g = leg_args(c, l);
            const bool big = leg_big(g.L);
            const unsigned nb = leg_grid(g.L, big, l == 0 ? LV_MAXB : 2048);
            if (l == 0) *nb_dots = (int)nb;
            LV_LAUNCH(leg_up_kernel(nu, l == 0, big), nb, g);
        }
        return 0;
    };
    if (c->l_vlds) {   // the complete V-cycle with levels >= l_lc in one LDS work-group
        const int Lc = c->l_lc;
        if (c->l_legs) TRY(legs_down(Lc));
        else
            for (int l = 0; l < Lc; ++l) TRY(down(l));
        hipLaunchKernelGGL(k_lv_vcoarse, dim3(1), dim3(1024), 0, c->stream, c->l_vargs);
        HIPC(hipGetLastError());
        if (c->l_legs) {
            int nbd = 0;
            TRY(legs_up(Lc, &nbd));
            return tail(nbd, 0);
        }
        for (int l = Lc - 1; l >= 1; --l) TRY(up(l));
        TRY(up0());
        return tail(nb0, 0);
    }
    if (c->l_fused) {
        const int Lc = c->l_lc;
        if (c->l_legs) TRY(legs_down(Lc));
        else
            for (int l = 0; l < Lc; ++l) TRY(down(l));
        CoarseNet cn;
        std::memset(&cn, 0, sizeof(cn));
        if (c->l_mode == LEARNED_NET) {
            const cfd_unet* u = c->l_net;
            cn.wsw = u->wsw;
            cn.w = u->w;
            cn.use_net = 1;
        }
        cn.vcycle = c->l_mode == LEARNED_VCYCLE ? 1 : 0;
        cn.gain = c->l_sigma / c->l_D;
        cn.one_a_one = c->l_1a1;
        hipLaunchKernelGGL(k_lv_coarse, dim3(1), dim3(1024), 0, c->stream, a, Lc, K, cn);
        HIPC(hipGetLastError());
        if (c->l_legs) {
            int nbd = 0;
            TRY(legs_up(Lc, &nbd));
            return tail(nbd, c->l_mode == LEARNED_NET ? 1 : 0);
        }
        for (int l = Lc - 1; l >= 1; --l) TRY(up(l));
        TRY(up0());
        return tail(nb0, c->l_mode == LEARNED_NET ? 1 : 0);
    }
    if (c->l_mode == LEARNED_VCYCLE) {   // the complete V-cycle, dense coarsest solve
        for (int l = 0; l < NL - 1; ++l) TRY(down(l));
        LV_LAUNCH(k_lv_dense, 1, a);
        for (int l = NL - 2; l >= 1; --l) TRY(up(l));
        if (NL >= 2) TRY(up0());
        else LV_LAUNCH512(k_fcg_dots, nb0, a, v);
        return tail(nb0, 0);
    }
    for (int l = 0; l < K; ++l) TRY(down(l));
    LV_LAUNCH(k_lv_cmode, nbK, a, K);
    if (c->l_mode != LEARNED_TEACHER) {
        const bool net = c->l_mode == LEARNED_NET;
        if (net) {
            LV_LAUNCH(k_lv_net_in, nimg, a, K, c->l_net->xin, c->l_H, c->l_W, nbK, c->l_1a1);
            // the network runs on this context's stream and idles once the solve is done
            cfd_unet* u = c->l_net;
            const hipStream_t s0 = u->stream;
            const int* k0 = u->skip;
            u->stream = c->stream;
            u->skip = &c->fcg->skip_net;
            const int rc = unet_forward(u, true);
            u->stream = s0;
            u->skip = k0;
            TRY(rc);
        }
        LV_LAUNCH(k_lv_net_out, cells(K), a, K, net ? (const float*)c->l_net->y : nullptr, c->l_W,
                  c->l_sigma / c->l_D, net ? 1 : 0, nbK, c->l_1a1);
        LV_LAUNCH(k_lv_smooth, cells(K), a, K, 1);
        LV_LAUNCH(k_lv_smooth, cells(K), a, K, 0);
    } else {
        const TeachVec tv = c->l_tv;
        LV_LAUNCH(k_lv_cmode_fin, 1, a, nbK, c->l_1a1, c->l_hcap);
        LV_LAUNCH(k_tch_edge, nbK, a, K, tv, 0);
        for (int it = 0; it < c->l_cycles; ++it) {
            for (int l = K; l < NL - 1; ++l) {
                LV_LAUNCH(k_lv_smooth, cells(l), a, l, 0);
                LV_LAUNCH(k_lv_restrict, cells(l + 1), a, l);
            }
            LV_LAUNCH(k_lv_dense, 1, a);
            for (int l = NL - 2; l >= K; --l) {
                LV_LAUNCH(k_lv_post, cells(l), a, l);
                LV_LAUNCH(k_lv_smooth, cells(l), a, l, 0);
            }
            LV_LAUNCH(k_tch_dots, nbK, a, K, tv, it);
            LV_LAUNCH(k_tch_dir, nbK, a, K, tv, it);
            LV_LAUNCH(k_tch_update, nbK, a, K, tv, it);
        }
        LV_LAUNCH(k_tch_edge, nbK, a, K, tv, 1);
        if (c->l_hcap > 0)
            LV_LAUNCH(k_lv_harvest, nimg, a, K, c->l_hin, c->l_hout, c->l_H, c->l_W, c->l_D);
    }
    for (int l = K - 1; l >= 1; --l) TRY(up(l));
    if (K >= 1) TRY(up0());
    else LV_LAUNCH512(k_fcg_dots, nb0, a, v);
    return tail(nb0, c->l_mode == LEARNED_NET ? 1 : 0);
}

void learned_drop_graph(cfd_ctx* c) {
    for (int k = 0; k < 3; ++k) {
        if (c->l_graph_x[k]) hipGraphExecDestroy(c->l_graph_x[k]);
        if (c->l_graph_g[k]) hipGraphDestroy(c->l_graph_g[k]);
        c->l_graph_x[k] = nullptr;
        c->l_graph_g[k] = nullptr;
    }
}

// the iteration with the current sweep count, captured once per count
int learned_capture(cfd_ctx* c) {
    const int k = c->l_sweeps - 2;
    HIPC(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    const int rc = learned_iteration(c);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) {
        if (g) hipGraphDestroy(g);
        return rc;
    }
    HIPC(e);
    c->l_graph_g[k] = g;
    HIPC(hipGraphInstantiate(&c->l_graph_x[k], g, nullptr, nullptr, 0));
    return 0;
}

// start and end of a learned projection as band launches (the steps of
// k_project before and after its solve, without grid barriers)
int learned_start(cfd_ctx* c, const ProjArgs& a) {
    const LvArgs la = lv_args(c);
    const int nb = std::min(c->nx, 512);
    // the extrapolated initial iterate (x0_extrapolate's) with the right-hand side
    int x0mode = 0;
    if (c->l_x0 && a.dt_override == 0.0) {
        x0mode = c->l_hist ? 2 : 1;
        c->l_hist = 1;
    }
    LV_LAUNCH512(k_lp_rhs, nb, a, c->geo, c->l_part, c->xprev, x0mode);
    LV_LAUNCH512(k_lp_res, nb, a, c->geo, la, nb);
    LV_LAUNCH(k_lp_init, 1, a, la, nb);
    return 0;
}

int learned_end(cfd_ctx* c, const ProjArgs& a) {
    const LvArgs la = lv_args(c);
    const int nb = std::min(c->nx, 512);
    LV_LAUNCH512(k_lp_xsum, nb, a, c->geo, c->l_part);
    if (a.correct) LV_LAUNCH512(k_lp_end<true>, nb, a, c->geo, la, nb);
    else LV_LAUNCH512(k_lp_end<false>, nb, a, c->geo, la, nb);
    LV_LAUNCH(k_lp_fin, 1, a, c->geo, c->l_part, nb, c->poll_d, ++c->poll_seq);
    if (a.zero_p_after) {
        const long n = (long)(c->nx + 2) * c->P;
        LV_LAUNCH(k_lp_pzero, (unsigned)((n + 255) / 256), a, n);
    }
    return 0;
}

// Wait for the poll record of the last end launches (k_lp_fin stores the
// solve's state, then the sequence number): the host spins on the pinned
// record instead of copying the state and synchronising the stream.  Every
// 1024 polls the stream is queried, so an error on the device is reported
// and a drained stream without the record is an error, not a hang.
int wait_poll(cfd_ctx* c) {
    const unsigned want = c->poll_seq;
    for (long spins = 1;; ++spins) {
        if (__atomic_load_n(&c->poll_h->seq, __ATOMIC_ACQUIRE) == want) return 0;
        if ((spins & 1023) == 0) {
            const hipError_t e = hipStreamQuery(c->stream);
            if (e == hipSuccess) {   // the record's stores precede the kernel's completion
                if (__atomic_load_n(&c->poll_h->seq, __ATOMIC_ACQUIRE) == want) return 0;
                return set_err(CFD_ERR_STATE, "the solve's state record did not arrive");
            }
            if (e != hipErrorNotReady) HIPC(e);
        }
        __builtin_ia32_pause();
    }
}

// A projection with the learned flexible CG: the start launches (right-hand
// side, initial residual), iterations, each batch followed by the end launches
// (pressure write-back, correction), which act once the device reports the
// solve finished; the host polls the state after each batch: first the
// iteration count of the last solve, then one iteration at a time.  A solve
// that does not converge is continued by the FP64 Jacobi-PCG of k_project when
// enabled (the end launches then leave the fields alone).  With adaptive
// sweeps, a solve after one of at least 4 iterations smooths with 3 sweeps per
// leg, and after 8 consecutive 3-sweep solves of at most 2 iterations with 2
// again.
int run_learned(cfd_ctx* c, ProjArgs a) {
    const int maxiter_all = a.maxiter;
    a.fcg = c->fcg;
    a.maxiter = std::min(a.maxiter, c->l_maxiter);
    if (c->l_sweeps_cfg < 0 && c->l_legs == 1) {
        // rare switches: consecutive solves with different preconditioners
        // leave uncorrelated errors in the two solutions the extrapolated
        // start is formed from
        c->l_low_run = (c->l_sweeps == 3 && c->l_prev_iters <= 2) ? c->l_low_run + 1 : 0;
        if (c->l_sweeps == 2 && c->l_prev_iters >= 4) c->l_sweeps = 3;
        else if (c->l_low_run >= 8) {
            c->l_sweeps = 2;
            c->l_low_run = 0;
        }
    }
    hipGraphExec_t gx = nullptr;
    if (c->l_graph) {
        if (!c->l_graph_x[c->l_sweeps - 2]) TRY(learned_capture(c));
        gx = c->l_graph_x[c->l_sweeps - 2];
    }
    const bool prof = c->prof_on;
    if (prof) {   // the whole projection is timed as one unit
        if (c->prof_used + 2 > c->prof_ev.size()) {
            for (int k = 0; k < 256; ++k) {
                hipEvent_t e;
                HIPC(hipEventCreate(&e));
                c->prof_ev.push_back(e);
            }
        }
        HIPC(hipEventRecord(c->prof_ev[c->prof_used], c->stream));
    }
    TRY(learned_start(c, a));
    int batch = std::max(1, c->l_prev_iters);
    int launched = 0;
    for (;;) {
        for (int q = 0; q < batch; ++q) {
            if (gx) HIPC(hipGraphLaunch(gx, c->stream));
            else TRY(learned_iteration(c));
        }
        launched += batch;
        TRY(learned_end(c, a));
        TRY(wait_poll(c));
        if (c->poll_h->s.done) break;
        if (launched > a.maxiter + 2)
            return set_err(CFD_ERR_STATE, "learned solver did not finish in %d iterations", launched);
        batch = 1;
    }
    const FcgState h = c->poll_h->s;
    c->l_solves += 1;
    c->l_iters_total += h.iters;
    if (h.info == 0) c->l_prev_iters = h.iters;
    if (h.info != 0 && c->l_fallback) {
        c->l_fallbacks += 1;
        ProjArgs f = a;
        f.method = PROJ_PCG64;
        f.maxiter = maxiter_all;
        TRY(launch_coop(c, f, false));
    } else if (h.info != 0) {
        // a failed solve without the fallback: its result as the solver left it
        a.end_always = 1;
        TRY(learned_end(c, a));
    }
    if (prof) {
        HIPC(hipEventRecord(c->prof_ev[c->prof_used + 1], c->stream));
        c->prof_used += 2;
    }
    return 0;
}

}  // namespace

#include "ens_host.hip"

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* cfd_last_error(void) { return g_err; }

int cfd_shutdown(void) {
    // every context, ensemble and network is destroyed by now (the caller's
    // duty); wait for the device so that no work is in flight when the
    // process's exit handlers tear the runtime down
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return 0;
    HIPC(hipDeviceSynchronize());
    return 0;
}

int cfd_device_count(int32_t* count) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return 0;
}

int cfd_ctx_create(const cfd_grid_desc* d, cfd_ctx** out) {
    if (!d || !out) return set_err(CFD_ERR_ARG, "null argument");
    if (d->nx < 4 || d->ny < 4) return set_err(CFD_ERR_ARG, "grid needs at least 4 cells per side");
    if (!d->fluid || !d->pcode || !d->pcomp || !d->diag_table || !d->inv_table)
        return set_err(CFD_ERR_ARG, "grid description lacks masks or tables");
    if (d->ncomp < 0 || d->ncomp > CFD_MAX_COMP)
        return set_err(CFD_ERR_LIMIT, "%d pure-Neumann pressure components (at most %d)", d->ncomp,
                       CFD_MAX_COMP);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return set_err(CFD_ERR_HIP, "no HIP device available");
    cfd_ctx* c = new cfd_ctx();
    c->nx = d->nx;
    c->ny = d->ny;
    c->dx = d->dx;
    c->dy = d->dy;
    c->periodic = d->periodic ? 1 : 0;
    c->P = ((d->ny + 2 + 31) / 32) * 32;
    c->ncomp = d->ncomp;
    for (int s = 0; s < 4; ++s) {
        c->kind[s] = d->bc_kind[s];
        c->vel[s] = d->bc_velocity[s];
    }
    c->geo.nx = c->nx;
    c->geo.ny = c->ny;
    c->geo.pu = c->geo.pv = c->geo.pc = c->P;
    c->geo.periodic = c->periodic;
    c->geo.dx = c->dx;
    c->geo.dy = c->dy;
    *out = nullptr;
    auto fail = [&](int rc) {
        cfd_ctx_destroy(c);
        return rc;
    };
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return fail(set_err(CFD_ERR_HIP, "cannot query the HIP device"));
    c->num_cu = prop.multiProcessorCount;
    c->device = dev;
    {   // set-up of the allocation-free entry point's staging on this device
        double* s = nullptr;
        const int rc = scratch_on(dev, &s);
        if (rc) return fail(rc);
    }
    if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess)
        return fail(set_err(CFD_ERR_HIP, "stream creation failed"));
    c->stream = c->own_stream;
    hipEventCreate(&c->ev0);
    hipEventCreate(&c->ev1);

    const size_t N = cells(c);
    int rc = 0;
    if ((rc = dalloc(c, &c->u, N)) || (rc = dalloc(c, &c->v, N)) || (rc = dalloc(c, &c->p, N)) ||
        (rc = dalloc(c, &c->F, N)) || (rc = dalloc(c, &c->G, N)) ||
        (rc = dalloc(c, &c->H[0][0], N)) || (rc = dalloc(c, &c->H[0][1], N)) ||
        (rc = dalloc(c, &c->H[1][0], N)) || (rc = dalloc(c, &c->H[1][1], N)) ||
        (rc = dalloc(c, &c->x, N)) || (rc = dalloc(c, &c->b, N)) || (rc = dalloc(c, &c->r, N)) ||
        (rc = dalloc(c, &c->q, N)) || (rc = dalloc(c, &c->d0, N)) || (rc = dalloc(c, &c->d1, N)) ||
        (rc = dalloc(c, &c->tmp0, N)) || (rc = dalloc(c, &c->tmp1, N)) ||
        (rc = dalloc(c, &c->rold, N)) ||
        (rc = dalloc(c, &c->r32, N)) || (rc = dalloc(c, &c->q32, N)) ||
        (rc = dalloc(c, &c->e0, N)) || (rc = dalloc(c, &c->e1, N)) ||
        (rc = dalloc(c, &c->fl, N)) || (rc = dalloc(c, &c->code, N)) ||
        (rc = dalloc(c, &c->comp, N)) || (rc = dalloc(c, &c->lut, 128)) ||
        (rc = dalloc(c, &c->comp_count, CFD_MAX_COMP)) || (rc = dalloc(c, &c->st, 1)) ||
        (rc = dalloc(c, &c->ring, RING)) || (rc = dalloc(c, &c->spin_err, 1)))
        return fail(rc);
    // persistent projection launch shape: one block per CU (or per row on small grids)
    c->proj_bs = std::min(PROJ_BS, ((c->ny + 63) / 64) * 64);
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_project, c->proj_bs, 0) !=
            hipSuccess || per_cu < 1)
        return fail(set_err(CFD_ERR_HIP, "projection kernel does not fit on a CU"));
    c->proj_blocks = std::max(1, std::min(c->num_cu, c->nx));
    c->red_blocks = std::max(1, std::min(c->num_cu, c->nx + 2));
    if ((rc = dalloc(c, &c->sync, (size_t)(c->proj_blocks + 2) * CFD_SLOT)) ||
        (rc = dalloc(c, &c->part, (size_t)REC_NV * c->red_blocks)) ||
        (rc = dalloc(c, &c->flags, (size_t)c->red_blocks)))
        return fail(rc);

    // masks and tables
    const int nx = c->nx, ny = c->ny, P = c->P;
    std::vector<uint8_t> flh(N, 0), codeh(N, 0), comph(N, 0);
    for (int i = 0; i < nx + 2; ++i)
        for (int j = 0; j < ny + 2; ++j) flh[(size_t)i * P + j] = d->fluid[(size_t)i * (ny + 2) + j] ? 1 : 0;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            codeh[(size_t)(i + 1) * P + j + 1] = d->pcode[(size_t)i * ny + j];
            const int cm = d->pcomp[(size_t)i * ny + j];
            if (cm > d->ncomp) return fail(set_err(CFD_ERR_ARG, "component label %d > ncomp", cm));
            comph[(size_t)(i + 1) * P + j + 1] = (uint8_t)cm;
        }
    c->flow.nfaces = 0.0;
    for (int i = 1; i <= nx; ++i)
        for (int j = 1; j <= ny; ++j)
            if (flh[(size_t)i * P + j] && flh[(size_t)(i + 1) * P + j]) c->flow.nfaces += 1.0;
    c->has_solid = 0;
    for (int i = 1; i <= nx && !c->has_solid; ++i)
        for (int j = 1; j <= ny; ++j)
            if (!flh[(size_t)i * P + j]) { c->has_solid = 1; break; }
    std::vector<double> luth(128, 0.0);
    for (int k = 0; k < 64; ++k) {
        luth[k] = d->diag_table[k];
        luth[64 + k] = d->inv_table[k];
    }
    luth[0] = 0.0;
    luth[64] = 0.0;
    std::vector<double> cc(CFD_MAX_COMP, 1.0);
    for (int k = 0; k < d->ncomp; ++k) cc[k] = d->comp_count[k];
    if (hipMemcpy(c->fl, flh.data(), N, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->code, codeh.data(), N, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->comp, comph.data(), N, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->lut, luth.data(), 128 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->comp_count, cc.data(), CFD_MAX_COMP * sizeof(double),
                  hipMemcpyHostToDevice) != hipSuccess)
        return fail(set_err(CFD_ERR_HIP, "mask upload failed"));
    for (int s = 0; s < 4; ++s) {
        const int len = s < 2 ? ny : nx;
        if ((rc = dalloc(c, &c->inflow[s], len)) || (rc = dalloc(c, &c->open[s], len)))
            return fail(rc);
        if (!d->open[s]) return fail(set_err(CFD_ERR_ARG, "open mask of side %d missing", s));
        if (hipMemcpy(c->open[s], d->open[s], len, hipMemcpyHostToDevice) != hipSuccess)
            return fail(set_err(CFD_ERR_HIP, "open mask upload failed"));
        if (d->inflow[s] && hipMemcpy(c->inflow[s], d->inflow[s], len * sizeof(double),
                                      hipMemcpyHostToDevice) != hipSuccess)
            return fail(set_err(CFD_ERR_HIP, "inflow upload failed"));
    }
    if (hipDeviceSynchronize() != hipSuccess) return fail(set_err(CFD_ERR_HIP, "setup sync failed"));
    *out = c;
    return 0;
}

int cfd_ctx_destroy(cfd_ctx* c) {
    if (!c) return 0;
    if (c->stream) hipStreamSynchronize(c->stream);
    for (hipEvent_t e : c->prof_ev) hipEventDestroy(e);
    learned_drop_graph(c);
    if (c->poll_h) hipHostFree(c->poll_h);
    for (void* p : c->allocs) hipFree(p);
    for (void* p : c->probes.mem) hipFree(p);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    if (c->own_stream) hipStreamDestroy(c->own_stream);
    delete c;
    return 0;
}

int cfd_ctx_set_stream(cfd_ctx* c, void* stream) {
    if (!c) return set_err(CFD_ERR_ARG, "null context");
    HIPC(hipStreamSynchronize(c->stream));
    c->stream = stream ? (hipStream_t)stream : c->own_stream;
    return 0;
}

int cfd_add_scalar(cfd_ctx* c, const cfd_scalar_desc* d, int32_t* id) {
    if (!c || !d || !id) return set_err(CFD_ERR_ARG, "null argument");
    if ((int)c->scalars.size() >= CFD_MAX_SCALARS)
        return set_err(CFD_ERR_LIMIT, "at most %d transported fields", CFD_MAX_SCALARS);
    if (d->limiter < 0 || d->limiter > 5) return set_err(CFD_ERR_ARG, "bad limiter %d", d->limiter);
    ScalarDev s;
    std::memset(&s, 0, sizeof(s));
    s.desc = *d;
    const size_t N = cells(c);
    TRY(dalloc(c, &s.phi[0], N));
    TRY(dalloc(c, &s.phi[1], N));
    for (int k = 0; k < 4; ++k) {
        const int len = k < 2 ? c->ny : c->nx;
        TRY(dalloc(c, &s.bt[k], len));
        TRY(dalloc(c, &s.bv[k], len));
        if (!d->bt[k] || !d->bv[k]) return set_err(CFD_ERR_ARG, "scalar BC arrays missing");
        for (int q = 0; q < len; ++q)
            if (d->bt[k][q] < 0 || d->bt[k][q] > 2)
                return set_err(CFD_ERR_ARG, "scalar BC code %lld", (long long)d->bt[k][q]);
        HIPC(hipMemcpy(s.bt[k], d->bt[k], len * sizeof(int64_t), hipMemcpyHostToDevice));
        HIPC(hipMemcpy(s.bv[k], d->bv[k], len * sizeof(double), hipMemcpyHostToDevice));
    }
    for (int k = 0; k < 4; ++k) s.desc.bt[k] = nullptr, s.desc.bv[k] = nullptr;
    c->scalars.push_back(s);
    *id = (int32_t)c->scalars.size() - 1;
    return 0;
}

int cfd_set_multigrid(cfd_ctx* c, const cfd_mg_desc* d) {
    if (!c || !d) return set_err(CFD_ERR_ARG, "null argument");
    if (d->nlev < 2 || d->nlev > MG_MAX_LEVELS) return set_err(CFD_ERR_ARG, "nlev %d", d->nlev);
    if (d->lev[0].nx != c->nx || d->lev[0].ny != c->ny)
        return set_err(CFD_ERR_ARG, "level 0 must be the context grid");
    const int last = d->nlev - 1;
    if (d->ncoarse != d->lev[last].nx * d->lev[last].ny || d->ncoarse > 4096 || !d->cinv)
        return set_err(CFD_ERR_ARG, "bad coarsest level (%d cells)", d->ncoarse);
    HIPC(hipStreamSynchronize(c->stream));
    MGDev h;
    std::memset(&h, 0, sizeof(h));
    h.nlev = d->nlev;
    h.omega = d->omega;
    h.wc = d->wc;
    h.ncoarse = d->ncoarse;
    for (int l = 0; l < d->nlev; ++l) {
        const cfd_mg_level_desc& L = d->lev[l];
        MGLevelDev& D = h.lev[l];
        D.nx = L.nx;
        D.ny = L.ny;
        if (l > 0) {
            const cfd_mg_level_desc& F = d->lev[l - 1];
            if (L.nx != (F.nx + 1) / 2 || L.ny != (F.ny + 1) / 2)
                return set_err(CFD_ERR_ARG, "level %d is not the 2x2 coarsening of level %d", l, l - 1);
            if (!L.diag || !L.ce || !L.cn) return set_err(CFD_ERR_ARG, "level %d coefficients", l);
        }
        D.P = l == 0 ? c->P : ((L.ny + 2 + 31) / 32) * 32;
        D.grid = l == 0 ? 1 : ((long)L.nx * L.ny > 4096 && h.lev[l - 1].grid ? 1 : 0);
        const size_t n = (size_t)(L.nx + 2) * D.P;
        float *b, *x, *t;
        TRY(dalloc(c, &b, n));
        TRY(dalloc(c, &x, n));
        TRY(dalloc(c, &t, n));
        D.b = b;
        D.x = x;
        D.t = t;
        if (l > 0) {
            float* co[3];
            const float* src[3] = {L.diag, L.ce, L.cn};
            for (int k = 0; k < 3; ++k) {
                TRY(dalloc(c, &co[k], n));
                HIPC(hipMemcpy2D(co[k] + D.P + 1, D.P * sizeof(float), src[k], L.ny * sizeof(float),
                                 L.ny * sizeof(float), L.nx, hipMemcpyHostToDevice));
            }
            D.diag = co[0];
            D.cE = co[1];
            D.cN = co[2];
        }
    }
    float* cinv;
    TRY(dalloc(c, &cinv, (size_t)d->ncoarse * d->ncoarse));
    HIPC(hipMemcpy(cinv, d->cinv, (size_t)d->ncoarse * d->ncoarse * sizeof(float),
                   hipMemcpyHostToDevice));
    h.cinv = cinv;
    if (!c->mg_dev) TRY(dalloc(c, &c->mg_dev, 1));
    HIPC(hipMemcpy(c->mg_dev, &h, sizeof(MGDev), hipMemcpyHostToDevice));
    c->mg_host = h;
    c->mg_ready = 1;
    c->l_ready = 0;   // a learned solver set up on the previous hierarchy is stale
    learned_drop_graph(c);
    return 0;
}

int cfd_set_learned(cfd_ctx* c, const cfd_learned_desc* d) {
    if (!c || !d) return set_err(CFD_ERR_ARG, "null argument");
    if (!c->mg_ready) return set_err(CFD_ERR_STATE, "cfd_set_multigrid must come first");
    const MGDev& mg = c->mg_host;
    if (d->level < 0 || d->level >= mg.nlev)
        return set_err(CFD_ERR_ARG, "level %d outside the hierarchy (0..%d)", d->level, mg.nlev - 1);
    if (d->mode < LEARNED_NET || d->mode > LEARNED_VCYCLE)
        return set_err(CFD_ERR_ARG, "mode %d", d->mode);
    const MGLevelDev& L = mg.lev[d->level];
    const int H = ((L.nx + 7) / 8) * 8, W = ((L.ny + 7) / 8) * 8;
    if (d->mode == LEARNED_NET) {
        if (!d->net) return set_err(CFD_ERR_ARG, "network mode without a network");
        if (d->net->H != H || d->net->W != W)
            return set_err(CFD_ERR_ARG, "network image %d x %d, level %d needs %d x %d", d->net->H,
                           d->net->W, d->level, H, W);
    }
    if (d->mode == LEARNED_TEACHER && (d->teacher_cycles < 1 || d->teacher_cycles > 64))
        return set_err(CFD_ERR_ARG, "teacher_cycles %d", d->teacher_cycles);
    if (d->harvest_cap < 0 || (d->harvest_cap > 0 && (!d->harvest_in || !d->harvest_out)))
        return set_err(CFD_ERR_ARG, "harvest buffers");
    if (d->max_iter < 1) return set_err(CFD_ERR_ARG, "max_iter %d", d->max_iter);
    if (d->sweeps != 0 && d->sweeps != -1 && (d->sweeps < 2 || d->sweeps > 4))
        return set_err(CFD_ERR_ARG, "smoothing sweeps %d (2, 3, 4 or -1)", d->sweeps);
    HIPC(hipStreamSynchronize(c->stream));
    if (!c->fcg) {
        TRY(dalloc(c, &c->fcg, 1));
        TRY(dalloc(c, &c->l_part, (size_t)LV_NSLOT * LV_SLOT));
        TRY(dalloc(c, &c->diag0f, 64));
        HIPC(hipHostMalloc((void**)&c->poll_h, sizeof(LpPoll), 0));
        std::memset(c->poll_h, 0, sizeof(LpPoll));
        HIPC(hipHostGetDevicePointer((void**)&c->poll_d, c->poll_h, 0));
        std::vector<double> lut(128);
        HIPC(hipMemcpy(lut.data(), c->lut, 128 * sizeof(double), hipMemcpyDeviceToHost));
        float df[64];
        for (int k = 0; k < 64; ++k) df[k] = (float)lut[k];
        HIPC(hipMemcpy(c->diag0f, df, sizeof(df), hipMemcpyHostToDevice));
    }
    FcgState h;
    std::memset(&h, 0, sizeof(h));
    h.done = 1;   // nothing to iterate before the first solve
    HIPC(hipMemcpy(c->fcg, &h, sizeof(h), hipMemcpyHostToDevice));
    const size_t n = (size_t)(L.nx + 2) * L.P;
    if (d->mode == LEARNED_TEACHER && c->l_tv_level != d->level) {
        float** v[7] = {&c->l_tv.e, &c->l_tv.r, &c->l_tv.ro, &c->l_tv.d0, &c->l_tv.d1, &c->l_tv.q,
                        &c->l_tv.bk};
        for (float** p : v) TRY(dalloc(c, p, n));
        c->l_tv_level = d->level;
    }
    learned_drop_graph(c);
    c->l_level = d->level;
    c->l_mode = d->mode;
    c->l_cycles = d->teacher_cycles;
    c->l_sigma = d->sigma;
    c->l_D = (float)(2.0 / (c->dx * c->dx) + 2.0 / (c->dy * c->dy));
    c->l_1a1 = d->one_a_one;
    c->l_fallback = d->fallback ? 1 : 0;
    c->l_graph = d->use_graph ? 1 : 0;
    c->l_maxiter = d->max_iter;
    c->l_x0 = d->x0_extrapolate ? 1 : 0;
    c->l_hist = 0;
    if (c->l_x0 && !c->xprev) TRY(dalloc(c, &c->xprev, (size_t)(c->nx + 2) * c->P));
    // the fused coarse end: a 32 x 32 level K below the fine level, entered at
    // the first level of at most 64 x 64 cells (never the fine level); one
    // work-group sweeps 64 x 64 cells faster than a launch costs, 128 x 128
    // cells slower
    c->l_fused = 0;
    c->l_vlds = 0;
    if (d->coarse_fused && d->mode == LEARNED_VCYCLE) {
        // the first level below the fine one with at most 64 x 64 cells, when
        // it and all coarser levels fit in the LDS pool of k_lv_vcoarse
        int lc = 1;
        while (lc < mg.nlev - 1 && (long)mg.lev[lc].nx * mg.lev[lc].ny > 64L * 64) ++lc;
        if (mg.nlev >= 2 && mg.nlev - lc <= VC_MAXL && mg.ncoarse <= VC_MAXC &&
            vc_need(mg, lc) <= VC_POOL) {
            c->l_lc = lc;
            c->l_vlds = 1;
            VcArgs& v = c->l_vargs;
            std::memset(&v, 0, sizeof(v));
            v.done = reinterpret_cast<const char*>(c->fcg) + offsetof(FcgState, done);
            v.done_stride = 0;
            v.cinv = mg.cinv;
            v.omega = mg.omega;
            v.wc = mg.wc;
            v.periodic = c->periodic;
            v.nl = mg.nlev - lc;
            v.nc = mg.ncoarse;
            for (int q = 0; q < v.nl; ++q) {
                const MGLevelDev& G = mg.lev[lc + q];
                VcLev& L = v.lev[q];
                L.nx = G.nx;
                L.ny = G.ny;
                L.pw = G.ny + 2;
                L.sz = (G.nx + 2) * L.pw;
                L.P = G.P;
                L.m = ((1ull << 32) + (unsigned long long)G.ny - 1) / (unsigned long long)G.ny;
                L.gd = G.diag;
                L.ge = G.cE;
                L.gn = G.cN;
                L.gb = G.b;
                L.gt = G.t;
                v.total += L.sz;
            }
            if (vc_assign(v) > VC_POOL) return set_err(CFD_ERR_STATE, "coarse-end pool");
            std::vector<float> img;
            if (vc_build_image(v, img)) return set_err(CFD_ERR_HIP, "coarse-end image");
            float* dimg = nullptr;
            TRY(dalloc(c, &dimg, img.size()));
            HIPC(hipMemcpy(dimg, img.data(), img.size() * sizeof(float), hipMemcpyHostToDevice));
            v.image = reinterpret_cast<const float4*>(dimg);
        }
    }
    if (!c->l_vlds && d->coarse_fused && d->mode != LEARNED_TEACHER && d->level >= 1 && L.nx == 32 &&
        L.ny == 32 && (d->mode != LEARNED_NET || (d->net->H == 32 && d->net->W == 32))) {
        int lc = d->level;
        while (lc - 1 >= 1 && (long)mg.lev[lc - 1].nx * mg.lev[lc - 1].ny <= 64L * 64) --lc;
        c->l_lc = lc;
        c->l_fused = 1;
    }
    c->l_legs = ((d->coarse_fused & 2) && (c->l_vlds || c->l_fused)) ? ((d->coarse_fused & 4) ? 2 : 1) : 0;
    c->l_pair = 0;
    if ((d->coarse_fused & 8) && c->l_vlds && c->l_legs == 1 && c->l_lc >= 3 && !c->periodic) {
        const MGLevelDev& F = mg.lev[c->l_lc - 2];
        const MGLevelDev& M = mg.lev[c->l_lc - 1];
        c->l_pair = F.nx == 2 * M.nx && F.ny == 2 * M.ny && !leg_big(F) && !leg_big(M) ? 1 : 0;
    }
    c->l_sweeps_cfg = d->sweeps < 0 ? -1 : d->sweeps > 2 ? d->sweeps : 2;
    c->l_sweeps = c->l_sweeps_cfg > 2 ? c->l_sweeps_cfg : 2;
    if (c->l_sweeps > 2 && c->l_legs != 1) {
        c->l_ready = 0;
        return set_err(CFD_ERR_ARG, "more than 2 smoothing sweeps need the fused legs (coarse_fused "
                                    "bits 0 and 1 without bit 2, a fused coarse end)");
    }
    c->l_H = H;
    c->l_W = W;
    c->l_net = d->mode == LEARNED_NET ? d->net : nullptr;
    c->l_hin = d->harvest_in;
    c->l_hout = d->harvest_out;
    c->l_hcap = d->harvest_cap;
    c->l_prev_iters = 0;
    c->l_low_run = 0;
    c->l_ready = 1;
    return 0;
}

int cfd_set_initial_iterate(cfd_ctx* c, int32_t mode) {
    if (!c) return set_err(CFD_ERR_ARG, "null argument");
    if (mode != 0 && mode != 1) return set_err(CFD_ERR_ARG, "initial iterate mode %d", mode);
    HIPC(hipStreamSynchronize(c->stream));
    if (mode == 1 && !c->xprev) TRY(dalloc(c, &c->xprev, (size_t)(c->nx + 2) * c->P));
    if (mode != c->l_x0) c->l_hist = 0;
    c->l_x0 = mode;
    return 0;
}

int cfd_initial_iterate(cfd_ctx* c, int32_t* mode, int32_t* has_history) {
    if (!c || !mode || !has_history) return set_err(CFD_ERR_ARG, "null argument");
    *mode = c->l_x0;
    *has_history = c->l_x0 ? c->l_hist : 0;
    return 0;
}

int cfd_learned_stats(cfd_ctx* c, cfd_learned_stats_t* out) {
    if (!c || !out) return set_err(CFD_ERR_ARG, "null argument");
    if (!c->fcg) return set_err(CFD_ERR_STATE, "no learned solver");
    HIPC(hipStreamSynchronize(c->stream));
    FcgState h;
    HIPC(hipMemcpy(&h, c->fcg, sizeof(h), hipMemcpyDeviceToHost));
    out->harvested = h.harvested;
    out->solves = c->l_solves;
    out->iterations = c->l_iters_total;
    out->fallbacks = c->l_fallbacks;
    out->last_iters = h.iters;
    out->last_info = h.info;
    out->last_relres = h.bnorm > 0.0 ? sqrt(h.rr) / h.bnorm : 0.0;
    out->net_off = h.net_off;
    out->last_sweeps = c->l_sweeps;
    return 0;
}

int cfd_learned_reset(cfd_ctx* c) {
    if (!c) return set_err(CFD_ERR_ARG, "null argument");
    if (!c->fcg) return set_err(CFD_ERR_STATE, "no learned solver");
    HIPC(hipStreamSynchronize(c->stream));
    FcgState h;
    HIPC(hipMemcpy(&h, c->fcg, sizeof(h), hipMemcpyDeviceToHost));
    h.harvested = 0;
    HIPC(hipMemcpy(c->fcg, &h, sizeof(h), hipMemcpyHostToDevice));
    c->l_solves = c->l_iters_total = c->l_fallbacks = 0;
    return 0;
}

int cfd_upload(cfd_ctx* c, int32_t field, const double* host) {
    if (!c || !host) return set_err(CFD_ERR_ARG, "null argument");
    int rows, cols, oi, oj;
    TRY(shape_of(c, field, &rows, &cols, &oi, &oj));
    int par = 0;
    TRY(host_parity(c, &par));
    if (field >= CFD_FIELD_NUT && field != CFD_FIELD_ALPHA && field != CFD_FIELD_XPREV)
        return set_err(CFD_ERR_ARG, "field %d is computed by a model", field);
    if (field == CFD_FIELD_ALPHA && !c->vof.on) return set_err(CFD_ERR_STATE, "no VOF model set");
    double* dst = field_ptr(c, field, par);
    if (field == CFD_FIELD_XPREV && !dst)
        return set_err(CFD_ERR_STATE, "no extrapolated initial iterate set");
    if (field == CFD_FIELD_X) c->l_hist = 0;       // a new iterate starts a new history
    if (field == CFD_FIELD_XPREV) c->l_hist = 1;   // ... which x_prev continues
    HIPC(hipMemcpy2DAsync(dst + (size_t)oi * c->P + oj, c->P * sizeof(double), host,
                          cols * sizeof(double), cols * sizeof(double), rows,
                          hipMemcpyHostToDevice, c->stream));
    if (field == CFD_FIELD_ALPHA) TRY(launch_vof_fill(c, 0, -1));   // vof.py:283-287
    HIPC(hipStreamSynchronize(c->stream));
    return 0;
}

int cfd_download(cfd_ctx* c, int32_t field, double* host) {
    if (!c || !host) return set_err(CFD_ERR_ARG, "null argument");
    int rows, cols, oi, oj;
    TRY(shape_of(c, field, &rows, &cols, &oi, &oj));
    int par = 0;
    TRY(host_parity(c, &par));
    const double* src = field_ptr(c, field, par);
    if (!src && field == CFD_FIELD_XPREV)
        return set_err(CFD_ERR_STATE, "no extrapolated initial iterate set");
    if (!src) return set_err(CFD_ERR_STATE, "field %d: no turbulence model set", field);
    HIPC(hipMemcpy2DAsync(host, cols * sizeof(double), src + (size_t)oi * c->P + oj,
                          c->P * sizeof(double), cols * sizeof(double), rows,
                          hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    return 0;
}

int cfd_upload_scalar(cfd_ctx* c, int32_t id, const double* host) {
    if (!c || !host || id < 0 || id >= (int)c->scalars.size())
        return set_err(CFD_ERR_ARG, "bad scalar id");
    int par = 0;
    TRY(host_parity(c, &par));
    ScalarDev& s = c->scalars[id];
    for (int k = 0; k < 2; ++k)
        HIPC(hipMemcpy2DAsync(s.phi[k], c->P * sizeof(double), host, (c->ny + 2) * sizeof(double),
                              (c->ny + 2) * sizeof(double), c->nx + 2, hipMemcpyHostToDevice,
                              c->stream));
    // fill_ghosts on the uploaded field (functions.py:366-369), result in phi[par]
    ScalarArgs a = scalar_args(c, s, par ^ 1);
    a.st = nullptr;
    hipLaunchKernelGGL(k_fill_ghosts, grid2d(c), dim3(BS2D), 0, c->stream, a, c->geo);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(s.phi[par ^ 1], s.phi[par], cells(c) * sizeof(double),
                        hipMemcpyDeviceToDevice, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    return 0;
}

int cfd_download_scalar(cfd_ctx* c, int32_t id, double* host) {
    if (!c || !host || id < 0 || id >= (int)c->scalars.size())
        return set_err(CFD_ERR_ARG, "bad scalar id");
    int par = 0;
    TRY(host_parity(c, &par));
    ScalarDev& s = c->scalars[id];
    HIPC(hipMemcpy2DAsync(host, (c->ny + 2) * sizeof(double), s.phi[par], c->P * sizeof(double),
                          (c->ny + 2) * sizeof(double), c->nx + 2, hipMemcpyDeviceToHost,
                          c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    return 0;
}

int cfd_set_state(cfd_ctx* c, const cfd_state* s) {
    if (!c || !s) return set_err(CFD_ERR_ARG, "null argument");
    DevState h;
    TRY(read_state(c, &h));
    const int old_par = (int)(h.n & 1), new_par = (int)(s->n & 1);
    if (old_par != new_par) {
        // keep the double-buffered fields where the new step count expects them
        for (int k = 0; k < 2; ++k)
            HIPC(hipMemcpyAsync(c->H[k][new_par ^ 1], c->H[k][old_par ^ 1],
                                cells(c) * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        for (auto& sc : c->scalars)
            HIPC(hipMemcpyAsync(sc.phi[new_par], sc.phi[old_par], cells(c) * sizeof(double),
                                hipMemcpyDeviceToDevice, c->stream));
    }
    h.t = s->t;
    h.dt = s->dt;
    h.dt_old = s->dt_old;
    h.gamma = s->gamma;
    h.flow_force = s->flow_force;
    h.n = s->n;
    h.has_old = s->has_old;
    h.done = s->done;
    h.error = s->error;
    HIPC(hipMemcpyAsync(c->st, &h, sizeof(DevState), hipMemcpyHostToDevice, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    return 0;
}

int cfd_get_state(cfd_ctx* c, cfd_state* s) {
    if (!c || !s) return set_err(CFD_ERR_ARG, "null argument");
    DevState h;
    TRY(read_state(c, &h));
    s->t = h.t;
    s->dt = h.dt;
    s->dt_old = h.dt_old;
    s->gamma = h.gamma;
    s->flow_force = h.flow_force;
    s->n = h.n;
    s->has_old = h.has_old;
    s->done = h.done;
    s->error = h.error;
    s->pad = 0;
    return 0;
}

int cfd_apply_bc(cfd_ctx* c) {
    if (!c) return set_err(CFD_ERR_ARG, "null context");
    TRY(launch_bc(c, false));
    return 0;
}

int cfd_project_initial(cfd_ctx* c, const cfd_step_params* prm) {
    if (!c || !prm) return set_err(CFD_ERR_ARG, "null argument");
    PredArgs pa = pred_args(c, prm, 0);
    hipLaunchKernelGGL(k_copy_predictor, grid2d(c), dim3(BS2D), 0, c->stream, pa, c->geo);
    HIPC(hipGetLastError());
    ProjArgs a = proj_args(c, prm->solver, prm->rtol, prm->maxiter);
    a.dt_override = 1.0;
    a.zero_p_after = 1;
    TRY(launch_project(c, a));
    c->l_hist = 0;   // no time history for an extrapolated initial iterate yet
    TRY(launch_bc(c, false));
    HIPC(hipStreamSynchronize(c->stream));
    return 0;
}

}  // extern "C"

namespace {

// the function objects of the step just issued (after k_record: st->n, t, dt
// are the new step's); ra.phi holds the scalars' current buffers
int launch_monitors(cfd_ctx* c, const RecordArgs& ra) {
    if (!c->probes.nsample && !c->favg.on) return 0;
    if (c->probes.nsample) {
        ProbeArgs a;
        std::memset(&a, 0, sizeof(a));
        a.u = c->u;
        a.v = c->v;
        a.p = c->p;
        a.fl = c->fl;
        for (int s = 0; s < ra.nscalar && s < MON_MAX_SCALARS; ++s) a.phi[s] = ra.phi[s];
        a.st = c->st;
        a.last_n = c->mon_last_n;
        a.nsample = c->probes.nsample;
        a.every = c->probes.every;
        a.cap = c->probes.cap;
        a.P = c->P;
        a.dx = c->geo.dx;
        a.dy = c->geo.dy;
        a.src = c->probes.src;
        a.base = c->probes.base;
        a.frac = c->probes.frac;
        a.times = c->probes.times;
        a.vals = c->probes.vals;
        a.count = c->probes.count;
        hipLaunchKernelGGL(k_probes, dim3(1), dim3(256), 0, c->stream, a);
        HIPC(hipGetLastError());
    }
    if (c->favg.on) {
        FavgArgs f;
        f.u = c->u;
        f.v = c->v;
        f.p = c->p;
        f.fl = c->fl;
        f.st = c->st;
        f.last_n = c->mon_last_n;
        f.t_start = c->favg.t_start;
        f.sum = c->favg.sum;
        f.sum2 = c->favg.sum2;
        f.T = c->favg.T;
        const long n = (long)c->nx * c->ny;
        hipLaunchKernelGGL(k_favg, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, f,
                           c->geo);
        HIPC(hipGetLastError());
    }
    hipLaunchKernelGGL(k_mon_fin, dim3(1), dim3(1), 0, c->stream, (const DevState*)c->st,
                       c->mon_last_n);
    HIPC(hipGetLastError());
    return 0;
}

}  // namespace

extern "C" {

int cfd_step(cfd_ctx* c, const cfd_step_params* prm, int32_t count) {
    if (!c || !prm) return set_err(CFD_ERR_ARG, "null argument");
    if (prm->nu <= 0.0) return set_err(CFD_ERR_ARG, "viscosity must be positive");
    int par = 0;
    TRY(host_parity(c, &par));
    StepParams sp;
    std::memset(&sp, 0, sizeof(sp));
    sp.nu = prm->nu;
    sp.tau = prm->tau;
    sp.scheme_ab2 = prm->scheme_ab2;
    sp.gamma_auto = prm->gamma_auto;
    sp.gamma_fixed = prm->gamma_fixed;
    sp.dt_fixed = prm->dt_fixed;
    sp.dt_model = prm->dt_model;
    sp.t_end = prm->t_end;
    sp.clamp_t_end = prm->clamp_t_end;
    sp.fx = prm->fx;
    sp.fy = prm->fy;
    if (prm->flow_control) {
        if (!(c->flow.nfaces > 0.0)) return set_err(CFD_ERR_ARG, "no fluid faces for flow control");
        if (!c->flow.u0) {
            const size_t N = cells(c);
            TRY(dalloc(c, &c->flow.u0, N));
            TRY(dalloc(c, &c->flow.v0, N));
            TRY(dalloc(c, &c->flow.pkeep, N));
            TRY(dalloc(c, &c->flow.ticket, 2));
            TRY(dalloc(c, &c->flow.acc, 1));
        }
    }
    HIPC(hipEventRecord(c->ev0, c->stream));
    for (int k = 0; k < count; ++k) {
        // steps that the device skips (run finished) keep the parity: the host
        // re-reads it only at the next synchronising call, so every step of a
        // batch is issued with the parity it will have if all steps run.
        const int pk = par ^ (k & 1);
        TRY(launch_bc(c, true));
        TurbDt td{nullptr, nullptr, nullptr, c->fl, nullptr};
        if (c->vof.on) {
            td.nut = c->vof.mu;
            td.rho = c->vof.rho;
            if (c->vof.sigma > 0.0) {
                VofArgs va = vof_args(c, c->vof.alpha[0], c->vof.alpha[0], -1);
                hipLaunchKernelGGL(k_vof_kappa, dim3((c->ny + 255) / 256, c->nx), dim3(256), 0,
                                   c->stream, va, c->geo);
                HIPC(hipGetLastError());
            }
        }
        if (c->turb.model != TURB_NONE) {
            TRY(launch_turb_pre(c, pk, prm->nu, true));
            td.nut = c->turb.nut;
            if (c->turb.model == TURB_SST) {
                td.dk = c->turb.dk;
                td.dw = c->turb.dw;
            }
        }
        if (prm->flow_control) {
            HIPC(hipMemcpyAsync(c->flow.u0, c->u, cells(c) * sizeof(double),
                                hipMemcpyDeviceToDevice, c->stream));
            HIPC(hipMemcpyAsync(c->flow.v0, c->v, cells(c) * sizeof(double),
                                hipMemcpyDeviceToDevice, c->stream));
        }
        hipLaunchKernelGGL(k_dt_gamma, dim3(c->red_blocks), dim3(RED_THREADS), 0,
                           c->stream, c->u, c->v, c->geo, sp, td, c->st, c->part, c->flags,
                           spin_guard(c));
        HIPC(hipGetLastError());
        PredArgs pa = pred_args(c, prm, pk);
        hipLaunchKernelGGL(k_predictor, grid2d(c), dim3(BS2D), 0, c->stream, pa, c->geo);
        HIPC(hipGetLastError());
        TRY(launch_project(c, proj_args(c, prm->solver, prm->rtol, prm->maxiter)));
        TRY(launch_bc(c, true));
        if (prm->flow_control) TRY(launch_flow_control(c, prm));
        for (int si = 0; si < (int)c->scalars.size(); ++si) {
            if (si == c->turb.sid_k || si == c->turb.sid_w) continue;   // SST below
            ScalarArgs a = scalar_args(c, c->scalars[si], pk);
            hipLaunchKernelGGL((k_scalar_transport<true, 1>), grid2d(c), dim3(BS2D), 0, c->stream,
                               a, c->geo);
            HIPC(hipGetLastError());
            hipLaunchKernelGGL(k_fill_ghosts, grid2d(c), dim3(BS2D), 0, c->stream, a, c->geo);
            HIPC(hipGetLastError());
        }
        if (c->turb.model == TURB_SST) TRY(launch_sst_post(c, pk));
        if (c->vof.on) TRY(launch_vof_post(c));
        RecordArgs ra;
        std::memset(&ra, 0, sizeof(ra));
        ra.u = c->u;
        ra.v = c->v;
        ra.fl = c->fl;
        ra.p = c->p;
        ra.has_solid = c->has_solid;
        ra.nscalar = (int)c->scalars.size();
        for (int s = 0; s < ra.nscalar; ++s) {
            ScalarDev& sd = c->scalars[s];
            ra.phi[s] = sd.phi[pk ^ 1];
            ra.diffs[s] = s == c->turb.sid_k ? c->turb.dk : (s == c->turb.sid_w ? c->turb.dw : nullptr);
            ra.D0[s] = sd.desc.diffusivity;
            ra.wallflux[s] = sd.desc.wall_flux;
            for (int k2 = 0; k2 < 4; ++k2) {
                ra.bt[s][k2] = sd.bt[k2];
                ra.bv[s][k2] = sd.bv[k2];
            }
        }
        ra.st = c->st;
        ra.part = c->part;
        ra.flags = c->flags;
        ra.spin = spin_guard(c);
        ra.ring = c->ring;
        ra.ring_len = RING;
        ra.scheme_ab2 = prm->scheme_ab2;
        ra.check_t_end = prm->check_t_end;
        ra.t_end = prm->t_end;
        ra.steady_tol = prm->steady_tol;
        ra.max_steps = prm->max_steps;
        {
            const dim3 gr(c->red_blocks), bl(RED_THREADS);
            switch (ra.nscalar) {
            case 0: hipLaunchKernelGGL(k_record<0>, gr, bl, 0, c->stream, ra, c->geo); break;
            case 1: hipLaunchKernelGGL(k_record<1>, gr, bl, 0, c->stream, ra, c->geo); break;
            case 2: hipLaunchKernelGGL(k_record<2>, gr, bl, 0, c->stream, ra, c->geo); break;
            case 3: hipLaunchKernelGGL(k_record<3>, gr, bl, 0, c->stream, ra, c->geo); break;
            default: hipLaunchKernelGGL(k_record<4>, gr, bl, 0, c->stream, ra, c->geo); break;
            }
        }
        HIPC(hipGetLastError());
        TRY(launch_monitors(c, ra));
    }
    HIPC(hipEventRecord(c->ev1, c->stream));
    return 0;
}

int cfd_sync(cfd_ctx* c) {
    if (!c) return set_err(CFD_ERR_ARG, "null context");
    HIPC(hipStreamSynchronize(c->stream));
    return check_spin(c);
}

int cfd_set_spin_limit(cfd_ctx* c, double seconds) {
    if (!c) return set_err(CFD_ERR_ARG, "null context");
    if (!(seconds == seconds) || seconds > 3600.0)
        return set_err(CFD_ERR_ARG, "spin limit %g s (at most 3600)", seconds);
    HIPC(hipStreamSynchronize(c->stream));
    c->spin_limit = seconds > 0.0 ? (long long)(seconds * 1e8) : 0;
    HIPC(hipMemsetAsync(c->spin_err, 0, sizeof(unsigned), c->stream));
    // flags that blocks raised after an abandoned gather would release the
    // next one early
    HIPC(hipMemsetAsync(c->flags, 0, (size_t)c->red_blocks * sizeof(unsigned), c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    return 0;
}

int cfd_last_elapsed_ms(cfd_ctx* c, double* ms) {
    if (!c || !ms) return set_err(CFD_ERR_ARG, "null argument");
    HIPC(hipEventSynchronize(c->ev1));
    float f = 0.0f;
    HIPC(hipEventElapsedTime(&f, c->ev0, c->ev1));
    *ms = f;
    return 0;
}

int cfd_set_profiling(cfd_ctx* c, int32_t on) {
    if (!c) return set_err(CFD_ERR_ARG, "null context");
    HIPC(hipStreamSynchronize(c->stream));
    c->prof_on = on ? 1 : 0;
    c->prof_used = 0;
    return 0;
}

int cfd_projection_times(cfd_ctx* c, double* total_ms, int64_t* launches) {
    if (!c || !total_ms || !launches) return set_err(CFD_ERR_ARG, "null argument");
    HIPC(hipStreamSynchronize(c->stream));
    double tot = 0.0;
    for (size_t k = 0; k + 1 < c->prof_used; k += 2) {
        float ms = 0.0f;
        HIPC(hipEventElapsedTime(&ms, c->prof_ev[k], c->prof_ev[k + 1]));
        tot += ms;
    }
    *total_ms = tot;
    *launches = (int64_t)(c->prof_used / 2);
    c->prof_used = 0;
    return 0;
}

int cfd_read_records(cfd_ctx* c, int64_t first_n, int32_t count, cfd_record* out) {
    if (!c || !out || count < 0) return set_err(CFD_ERR_ARG, "bad arguments");
    static_assert(sizeof(cfd_record) == sizeof(StepRecord), "record layout");
    DevState h;
    TRY(read_state(c, &h));
    if (first_n < 1 || first_n + count - 1 > h.n || h.n - first_n >= RING)
        return set_err(CFD_ERR_ARG, "records %lld..%lld not available (step %lld, ring %d)",
                       (long long)first_n, (long long)(first_n + count - 1), (long long)h.n, RING);
    for (int k = 0; k < count; ++k) {
        const long long n = first_n + k;
        HIPC(hipMemcpyAsync(&out[k], &c->ring[(n - 1) % RING], sizeof(StepRecord),
                            hipMemcpyDeviceToHost, c->stream));
    }
    HIPC(hipStreamSynchronize(c->stream));
    return 0;
}

int cfd_pressure_solve(cfd_ctx* c, const double* rhs, double* p_out, int32_t solver, double rtol,
                       int32_t maxiter, int32_t* iters, double* relres) {
    if (!c || !rhs || !p_out) return set_err(CFD_ERR_ARG, "null argument");
    HIPC(hipMemcpy2DAsync(c->tmp0 + c->P + 1, c->P * sizeof(double), rhs, c->ny * sizeof(double),
                          c->ny * sizeof(double), c->nx, hipMemcpyHostToDevice, c->stream));
    DevState h;
    TRY(read_state(c, &h));
    const int done = h.done;
    if (done) {
        h.done = 0;
        HIPC(hipMemcpyAsync(c->st, &h, sizeof(DevState), hipMemcpyHostToDevice, c->stream));
    }
    ProjArgs a = proj_args(c, solver, rtol, maxiter);
    a.rhs_in = c->tmp0;
    a.p = c->tmp1;
    a.correct = 0;
    a.dt_override = 1.0;
    TRY(launch_project(c, a));
    HIPC(hipMemcpy2DAsync(p_out, c->ny * sizeof(double), c->tmp1 + c->P + 1,
                          c->P * sizeof(double), c->ny * sizeof(double), c->nx,
                          hipMemcpyDeviceToHost, c->stream));
    DevState h2;
    TRY(read_state(c, &h2));
    if (iters) *iters = h2.cg_iters;
    if (relres) *relres = h2.cg_relres;
    if (done) {
        h2.done = 1;
        HIPC(hipMemcpyAsync(c->st, &h2, sizeof(DevState), hipMemcpyHostToDevice, c->stream));
        HIPC(hipStreamSynchronize(c->stream));
    }
    return h2.cg_info > 0 ? h2.cg_info : 0;
}

// ---- device memory helpers ----
int cfd_dev_alloc(int64_t bytes, void** out) {
    if (!out || bytes < 0) return set_err(CFD_ERR_ARG, "bad arguments");
    *out = nullptr;
    HIPC(hipMalloc(out, std::max<int64_t>(bytes, 1)));
    return 0;
}

int cfd_dev_free(void* ptr) {
    if (ptr) HIPC(hipFree(ptr));
    return 0;
}

int cfd_memcpy(void* dst, const void* src, int64_t bytes, int32_t kind) {
    if ((!dst || !src) && bytes > 0) return set_err(CFD_ERR_ARG, "null pointer");
    hipMemcpyKind k;
    switch (kind) {
    case 1: k = hipMemcpyHostToDevice; break;
    case 2: k = hipMemcpyDeviceToHost; break;
    case 3: k = hipMemcpyDeviceToDevice; break;
    default: return set_err(CFD_ERR_ARG, "bad copy kind %d", kind);
    }
    if (bytes > 0) HIPC(hipMemcpy(dst, src, (size_t)bytes, k));
    HIPC(hipDeviceSynchronize());
    return 0;
}

// ---- stateless kernels on reference-layout device arrays ----
int cfd_momentum_tendencies(const double* u, const double* v, const uint8_t* fluid, int32_t nx,
                            int32_t ny, double dx, double dy, double nu0, double gamma,
                            int32_t periodic, double* hu, double* hv, void* stream) {
    if (!u || !v || !fluid || !hu || !hv) return set_err(CFD_ERR_ARG, "null array");
    if (nx < 4 || ny < 4) return set_err(CFD_ERR_ARG, "grid too small");
    Geo g{nx, ny, ny + 2, ny + 1, ny + 2, periodic ? 1 : 0, dx, dy};
    hipLaunchKernelGGL(k_momentum_tendencies<0>, dim3((ny + 2 + BS2D - 1) / BS2D, nx + 2),
                       dim3(BS2D), 0, (hipStream_t)stream, u, v, fluid, g, nu0,
                       (const double*)nullptr, gamma, hu, hv);
    HIPC(hipGetLastError());
    return 0;
}

int cfd_momentum_tendencies_var(const double* u, const double* v, const uint8_t* fluid,
                                int32_t nx, int32_t ny, double dx, double dy, const double* nu,
                                double gamma, int32_t periodic, double* hu, double* hv,
                                void* stream) {
    if (!u || !v || !fluid || !nu || !hu || !hv) return set_err(CFD_ERR_ARG, "null array");
    if (nx < 4 || ny < 4) return set_err(CFD_ERR_ARG, "grid too small");
    Geo g{nx, ny, ny + 2, ny + 1, ny + 2, periodic ? 1 : 0, dx, dy};
    hipLaunchKernelGGL(k_momentum_tendencies<1>, dim3((ny + 2 + BS2D - 1) / BS2D, nx + 2),
                       dim3(BS2D), 0, (hipStream_t)stream, u, v, fluid, g, 0.0, nu, gamma, hu, hv);
    HIPC(hipGetLastError());
    return 0;
}

int cfd_obstacle_bc(double* u, double* v, const uint8_t* fluid, int32_t nx, int32_t ny,
                    int32_t periodic, void* stream) {
    if (!u || !v || !fluid) return set_err(CFD_ERR_ARG, "null array");
    if (nx < 4 || ny < 4) return set_err(CFD_ERR_ARG, "grid too small");
    Geo g{nx, ny, ny + 2, ny + 1, ny + 2, periodic ? 1 : 0, 1.0, 1.0};
    hipLaunchKernelGGL(k_obstacle_bc, dim3((ny + 2 + BS2D - 1) / BS2D, nx + 2), dim3(BS2D), 0,
                       (hipStream_t)stream, u, v, fluid, g, (const DevState*)nullptr);
    HIPC(hipGetLastError());
    return 0;
}

int cfd_scalar_transport(const double* phi, const double* u, const double* v,
                         const uint8_t* fluid, int32_t nx, int32_t ny, double dx, double dy,
                         const double* diff, double d0, int32_t limiter, int32_t periodic,
                         const int64_t* bt_l, const double* bv_l, const int64_t* bt_r,
                         const double* bv_r, const int64_t* bt_b, const double* bv_b,
                         const int64_t* bt_t, const double* bv_t, int32_t obs_t, double obs_v,
                         double* out, void* stream) {
    if (!phi || !u || !v || !fluid || !out || !bt_l || !bt_r || !bt_b || !bt_t || !bv_l ||
        !bv_r || !bv_b || !bv_t)
        return set_err(CFD_ERR_ARG, "null array");
    if (nx < 4 || ny < 4) return set_err(CFD_ERR_ARG, "grid too small");
    if (limiter < 0 || limiter > 5) return set_err(CFD_ERR_ARG, "bad limiter %d", limiter);
    Geo g{nx, ny, ny + 2, ny + 1, ny + 2, periodic ? 1 : 0, dx, dy};
    ScalarArgs a;
    std::memset(&a, 0, sizeof(a));
    a.phi_in = phi;
    a.out = out;
    a.u = u;
    a.v = v;
    a.fl = fluid;
    a.diff = diff;
    a.D0 = d0;
    a.limiter = limiter;
    a.obs_t = obs_t;
    a.obs_v = obs_v;
    a.bt[0] = bt_l; a.bt[1] = bt_r; a.bt[2] = bt_b; a.bt[3] = bt_t;
    a.bv[0] = bv_l; a.bv[1] = bv_r; a.bv[2] = bv_b; a.bv[3] = bv_t;
    const dim3 grid((ny + 2 + BS2D - 1) / BS2D, nx + 2);
    if (diff)
        hipLaunchKernelGGL((k_scalar_transport<false, 0>), grid, dim3(BS2D), 0,
                           (hipStream_t)stream, a, g);
    else
        hipLaunchKernelGGL((k_scalar_transport<true, 0>), grid, dim3(BS2D), 0,
                           (hipStream_t)stream, a, g);
    HIPC(hipGetLastError());
    return 0;
}

// ---- function objects on the device --------------------------------------
int cfd_probes_set(cfd_ctx* c, const cfd_probe_desc* d) {
    if (!c || !d) return set_err(CFD_ERR_ARG, "null argument");
    HIPC(hipStreamSynchronize(c->stream));
    auto release = [&]() {
        for (void* q : c->probes.mem) hipFree(q);
        c->probes.mem.clear();
        c->probes.nsample = 0;
    };
    if (d->nsample == 0) {
        release();
        return 0;
    }
    if (d->nsample < 0 || d->every < 1 || d->capacity < 1 || !d->src || !d->corner || !d->frac)
        return set_err(CFD_ERR_ARG, "probe descriptor");
    const int n = d->nsample;
    const int nsc = std::min((int)c->scalars.size(), MON_MAX_SCALARS);
    std::vector<long long> base(n);
    for (int s = 0; s < n; ++s) {
        const int sr = d->src[s];
        if (sr < CFD_PROBE_U || sr >= CFD_PROBE_SCALAR + nsc)
            return set_err(CFD_ERR_ARG, "probe %d: field %d", s, sr);
        // host-layout extent of the field the corner indexes; the
        // interpolation cell's upper corner (i+1, j+1) must lie inside it
        const int rows = sr == CFD_PROBE_U ? c->nx + 1 : (sr == CFD_PROBE_V ? c->nx + 2 : c->nx);
        const int cols = sr == CFD_PROBE_U ? c->ny + 2 : (sr == CFD_PROBE_V ? c->ny + 1 : c->ny);
        const int off = sr >= CFD_PROBE_PRESSURE ? 1 : 0;   // cells sit inside the ghost ring
        const int i = d->corner[2 * s], j = d->corner[2 * s + 1];
        if (i < 0 || i + 1 >= rows || j < 0 || j + 1 >= cols)
            return set_err(CFD_ERR_ARG, "probe %d: corner (%d, %d) needs %d x %d with room for "
                           "(i+1, j+1)", s, i, j, rows, cols);
        base[s] = (long long)(i + off) * c->P + (j + off);
    }
    release();   // a rejected descriptor leaves the previous probes in place
    auto alloc = [&](void** q, size_t bytes) -> int {
        HIPC(hipMalloc(q, std::max<size_t>(bytes, 8)));
        HIPC(hipMemsetAsync(*q, 0, std::max<size_t>(bytes, 8), c->stream));
        HIPC(hipStreamSynchronize(c->stream));
        c->probes.mem.push_back(*q);
        return 0;
    };
    TRY(alloc((void**)&c->probes.src, sizeof(int) * n));
    TRY(alloc((void**)&c->probes.base, sizeof(long long) * n));
    TRY(alloc((void**)&c->probes.frac, sizeof(double) * 2L * n));
    TRY(alloc((void**)&c->probes.times, sizeof(double) * d->capacity));
    TRY(alloc((void**)&c->probes.vals, sizeof(double) * (size_t)d->capacity * n));
    TRY(alloc((void**)&c->probes.count, sizeof(int) * 2));
    if (!c->mon_last_n) TRY(dalloc(c, &c->mon_last_n, 1));
    HIPC(hipMemcpy(c->probes.src, d->src, sizeof(int) * n, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(c->probes.base, base.data(), sizeof(long long) * n, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(c->probes.frac, d->frac, sizeof(double) * 2L * n, hipMemcpyHostToDevice));
    DevState h;
    TRY(read_state(c, &h));
    HIPC(hipMemcpy(c->mon_last_n, &h.n, sizeof(long long), hipMemcpyHostToDevice));
    c->probes.every = d->every;
    c->probes.cap = d->capacity;
    c->probes.nsample = n;
    return 0;
}

int cfd_probes_read(cfd_ctx* c, int32_t max_records, double* times, double* values,
                    int32_t* nrecords, int64_t* lost) {
    if (!c || !nrecords) return set_err(CFD_ERR_ARG, "null argument");
    *nrecords = 0;
    if (lost) *lost = 0;
    if (!c->probes.nsample) return 0;
    int cnt[2];
    HIPC(hipMemcpyAsync(cnt, c->probes.count, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    const int n = std::min(cnt[0], std::max(0, max_records));
    if (n < cnt[0]) return set_err(CFD_ERR_LIMIT, "%d probe records held, room for %d", cnt[0],
                                   max_records);
    if (n > 0) {
        if (!times || !values) return set_err(CFD_ERR_ARG, "null output");
        HIPC(hipMemcpy(times, c->probes.times, sizeof(double) * n, hipMemcpyDeviceToHost));
        HIPC(hipMemcpy(values, c->probes.vals, sizeof(double) * (size_t)n * c->probes.nsample,
                       hipMemcpyDeviceToHost));
    }
    HIPC(hipMemsetAsync(c->probes.count, 0, sizeof(int) * 2, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    *nrecords = n;
    if (lost) *lost = cnt[1];
    return 0;
}

int cfd_favg_set(cfd_ctx* c, int32_t enable, double t_start) {
    if (!c) return set_err(CFD_ERR_ARG, "null context");
    HIPC(hipStreamSynchronize(c->stream));
    c->favg.on = 0;
    if (!enable) return 0;
    const size_t n = 3 * (size_t)c->nx * c->ny;
    if (!c->favg.sum) {
        TRY(dalloc(c, &c->favg.sum, n));
        TRY(dalloc(c, &c->favg.sum2, n));
        TRY(dalloc(c, &c->favg.T, 1));
    }
    // on the context stream, ordered before the accumulating kernels
    HIPC(hipMemsetAsync(c->favg.sum, 0, n * sizeof(double), c->stream));
    HIPC(hipMemsetAsync(c->favg.sum2, 0, n * sizeof(double), c->stream));
    HIPC(hipMemsetAsync(c->favg.T, 0, sizeof(double), c->stream));
    if (!c->mon_last_n) TRY(dalloc(c, &c->mon_last_n, 1));
    DevState h;
    TRY(read_state(c, &h));
    HIPC(hipMemcpy(c->mon_last_n, &h.n, sizeof(long long), hipMemcpyHostToDevice));
    c->favg.t_start = t_start;
    c->favg.on = 1;
    return 0;
}

int cfd_favg_read(cfd_ctx* c, double* T, double* sum, double* sum2) {
    if (!c || !T) return set_err(CFD_ERR_ARG, "null argument");
    if (!c->favg.sum) return set_err(CFD_ERR_STATE, "field averaging was never enabled");
    HIPC(hipStreamSynchronize(c->stream));
    const size_t n = 3 * (size_t)c->nx * c->ny;
    HIPC(hipMemcpy(T, c->favg.T, sizeof(double), hipMemcpyDeviceToHost));
    if (sum) HIPC(hipMemcpy(sum, c->favg.sum, n * sizeof(double), hipMemcpyDeviceToHost));
    if (sum2) HIPC(hipMemcpy(sum2, c->favg.sum2, n * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

}  // extern "C"

namespace {

// the node field `kind` of the current velocity into dst, (nx+1) x (ny+1),
// on the context's stream
int launch_derived(cfd_ctx* c, int32_t kind, double* dst) {
    const size_t n = (size_t)(c->nx + 1) * (c->ny + 1);
    if (kind == CFD_DERIVED_VORTICITY) {
        hipLaunchKernelGGL(k_vort_nodes, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                           (const double*)c->u, (const double*)c->v, c->geo, c->P, dst);
    } else {
        hipLaunchKernelGGL(k_psi_bottom, dim3(1), dim3(1024), 0, c->stream, (const double*)c->v,
                           (const uint8_t*)c->fl, c->geo, c->P, dst);
        HIPC(hipGetLastError());
        hipLaunchKernelGGL(k_psi_cols, dim3((unsigned)((c->nx + 1 + 255) / 256)), dim3(256), 0,
                           c->stream, (const double*)c->u, (const uint8_t*)c->fl, c->geo, c->P,
                           dst);
    }
    HIPC(hipGetLastError());
    return 0;
}

}  // namespace

extern "C" {

int cfd_download_derived(cfd_ctx* c, int32_t kind, double* host) {
    if (!c || !host) return set_err(CFD_ERR_ARG, "null argument");
    if (kind != CFD_DERIVED_VORTICITY && kind != CFD_DERIVED_STREAMFUNCTION)
        return set_err(CFD_ERR_ARG, "derived field %d", kind);
    const size_t n = (size_t)(c->nx + 1) * (c->ny + 1);
    if (!c->derived) TRY(dalloc(c, &c->derived, n));
    TRY(launch_derived(c, kind, c->derived));
    HIPC(hipMemcpyAsync(host, c->derived, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    return 0;
}

int cfd_download_derived_noalloc(cfd_ctx* c, int32_t kind, double* host) {
    if (!c || !host) return set_err(CFD_ERR_ARG, "null argument");
    if (kind != CFD_DERIVED_VORTICITY && kind != CFD_DERIVED_STREAMFUNCTION)
        return set_err(CFD_ERR_ARG, "derived field %d", kind);
    const long n = (long)(c->nx + 1) * (c->ny + 1);
    if (n > CFD_DERIVED_NOALLOC_MAX)
        return set_err(CFD_ERR_LIMIT, "%d x %d nodes exceed the allocation-free staging of %ld "
                       "nodes (use cfd_download_derived)", c->nx + 1, c->ny + 1,
                       (long)CFD_DERIVED_NOALLOC_MAX);
    // this device's staging (set up when the context was created); both
    // kernels write every one of the n nodes, so nothing needs clearing
    double* s = nullptr;
    TRY(scratch_on(c->device, &s));
    TRY(launch_derived(c, kind, s));
    HIPC(hipMemcpyAsync(host, s, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    return 0;
}

// ---- turbulence models (turbulence.py) -------------------------------------
int cfd_set_turbulence(cfd_ctx* c, const cfd_turb_desc* d) {
    if (!c || !d) return set_err(CFD_ERR_ARG, "null argument");
    if (d->model != CFD_TURB_SMAGORINSKY && d->model != CFD_TURB_SST)
        return set_err(CFD_ERR_ARG, "turbulence model %d", d->model);
    if (c->turb.model != TURB_NONE) return set_err(CFD_ERR_STATE, "a turbulence model is set");
    if (!(d->nu > 0.0)) return set_err(CFD_ERR_ARG, "viscosity must be positive");
    const int nx = c->nx, ny = c->ny, P = c->P;
    const int nsc = (int)c->scalars.size();
    if (d->model == CFD_TURB_SMAGORINSKY && !d->l2) return set_err(CFD_ERR_ARG, "l2 missing");
    if (d->model == CFD_TURB_SST) {
        if (!d->wall_distance || !d->wall_adjacent || !d->omega_wall)
            return set_err(CFD_ERR_ARG, "SST wall arrays missing");
        if (d->sid_k < 0 || d->sid_k >= nsc || d->sid_omega < 0 || d->sid_omega >= nsc ||
            d->sid_k == d->sid_omega)
            return set_err(CFD_ERR_ARG, "SST scalars %d, %d (of %d)", d->sid_k, d->sid_omega, nsc);
    }
    HIPC(hipStreamSynchronize(c->stream));
    const size_t N = cells(c);
    // inner nx x ny host arrays into the padded device layout
    auto put = [&](double** dst, const double* src) -> int {
        TRY(dalloc(c, dst, N));
        std::vector<double> h(N, 0.0);
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j) h[(size_t)(i + 1) * P + j + 1] = src[(size_t)i * ny + j];
        HIPC(hipMemcpy(*dst, h.data(), N * sizeof(double), hipMemcpyHostToDevice));
        return 0;
    };
    TRY(dalloc(c, &c->turb.nut, N));
    if (d->model == CFD_TURB_SMAGORINSKY) {
        TRY(put(&c->turb.l2, d->l2));
    } else {
        TRY(put(&c->turb.d, d->wall_distance));
        TRY(put(&c->turb.wwall, d->omega_wall));
        TRY(dalloc(c, &c->turb.wadj, N));
        std::vector<uint8_t> h(N, 0);
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j)
                h[(size_t)(i + 1) * P + j + 1] = d->wall_adjacent[(size_t)i * ny + j] ? 1 : 0;
        HIPC(hipMemcpy(c->turb.wadj, h.data(), N, hipMemcpyHostToDevice));
        TRY(dalloc(c, &c->turb.F1, N));
        TRY(dalloc(c, &c->turb.cross, N));
        TRY(dalloc(c, &c->turb.dk, N));
        TRY(dalloc(c, &c->turb.dw, N));
        c->turb.sid_k = d->sid_k;
        c->turb.sid_w = d->sid_omega;
    }
    c->turb.model = d->model == CFD_TURB_SMAGORINSKY ? TURB_SMAGORINSKY : TURB_SST;
    c->turb.nu = d->nu;
    // the model's first evaluation at attach (turbulence.py:144, 239)
    int par = 0;
    TRY(host_parity(c, &par));
    TRY(launch_turb_pre(c, par, d->nu, false));
    HIPC(hipStreamSynchronize(c->stream));
    return 0;
}

// ---- two-phase VOF (vof.py) ---------------------------------------------------
int cfd_set_vof(cfd_ctx* c, const cfd_vof_desc* d) {
    if (!c || !d || !d->alpha0) return set_err(CFD_ERR_ARG, "null argument");
    if (c->vof.on) return set_err(CFD_ERR_STATE, "a VOF model is set");
    if (c->turb.model != TURB_NONE)
        return set_err(CFD_ERR_STATE, "VOF and a turbulence model together are not supported");
    if (!(d->rho1 > 0.0 && d->rho2 > 0.0) || !(d->cfl_alpha > 0.0))
        return set_err(CFD_ERR_ARG, "densities and cfl_alpha must be positive");
    if (d->npinned < 0 || d->npinned > CFD_MAX_COMP || (d->npinned && !d->pinned))
        return set_err(CFD_ERR_ARG, "pinned cells");
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_project_var, c->proj_bs, 0) !=
            hipSuccess || per_cu < 1)
        return set_err(CFD_ERR_HIP, "variable-density projection kernel does not fit on a CU");
    HIPC(hipStreamSynchronize(c->stream));
    const int nx = c->nx, ny = c->ny, P = c->P;
    const size_t N = cells(c);
    TRY(dalloc(c, &c->vof.alpha[0], N));
    TRY(dalloc(c, &c->vof.alpha[1], N));
    TRY(dalloc(c, &c->vof.rho, N));
    TRY(dalloc(c, &c->vof.mu, N));
    TRY(dalloc(c, &c->vof.kappa, N));
    TRY(dalloc(c, &c->vof.cE, N));
    TRY(dalloc(c, &c->vof.cN, N));
    TRY(dalloc(c, &c->vof.diag, N));
    std::vector<double> h(N, 0.0);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) h[(size_t)(i + 1) * P + j + 1] = d->alpha0[(size_t)i * ny + j];
    HIPC(hipMemcpy(c->vof.alpha[0], h.data(), N * sizeof(double), hipMemcpyHostToDevice));
    std::vector<int> pin(std::max(1, (int)d->npinned), -1);
    for (int k = 0; k < d->npinned; ++k) {
        const int pi = d->pinned[2 * k], pj = d->pinned[2 * k + 1];
        if (pi < 0 || pi >= nx || pj < 0 || pj >= ny) return set_err(CFD_ERR_ARG, "pinned cell");
        pin[k] = (pi + 1) * P + pj + 1;
    }
    TRY(dalloc(c, &c->vof.pinned, pin.size()));
    TRY(dalloc(c, &c->vof.part, (size_t)3 * c->red_blocks));
    TRY(dalloc(c, &c->vof.ticket, 2));
    TRY(dalloc(c, &c->vof.acc, 2));
    HIPC(hipMemcpy(c->vof.pinned, pin.data(), pin.size() * sizeof(int), hipMemcpyHostToDevice));
    c->vof.npinned = d->npinned;
    // multigrid levels: 2x2 aggregation (2x1 once a side is one cell) down to
    // at most VMG_COARSE cells; a grid too large for VMG_MAX levels keeps the
    // Jacobi preconditioner
    std::memset(&c->vof.mg, 0, sizeof(c->vof.mg));
    {
        VMG& mg = c->vof.mg;
        mg.lev[0].nx = nx;
        mg.lev[0].ny = ny;
        mg.lev[0].P = P;
        TRY(dalloc(c, &mg.lev[0].x, N));
        TRY(dalloc(c, &mg.lev[0].t, N));
        int L = 1;
        while (L < VMG_MAX && (long)mg.lev[L - 1].nx * mg.lev[L - 1].ny > VMG_COARSE) {
            VLev& V = mg.lev[L];
            V.nx = (mg.lev[L - 1].nx + 1) / 2;
            V.ny = (mg.lev[L - 1].ny + 1) / 2;
            V.P = V.ny + 2;
            const size_t M = (size_t)(V.nx + 2) * V.P;
            TRY(dalloc(c, &V.cE, M));
            TRY(dalloc(c, &V.cN, M));
            TRY(dalloc(c, &V.diag, M));
            TRY(dalloc(c, &V.x, M));
            TRY(dalloc(c, &V.t, M));
            TRY(dalloc(c, &V.b, M));
            ++L;
        }
        const bool fits = (long)mg.lev[L - 1].nx * mg.lev[L - 1].ny <= VMG_COARSE;
        mg.nlev = (L >= 2 && fits && !d->jacobi_only) ? L : 0;
    }
    c->vof.rho1 = d->rho1;
    c->vof.rho2 = d->rho2;
    c->vof.mu1 = d->mu1;
    c->vof.mu2 = d->mu2;
    c->vof.sigma = d->sigma;
    c->vof.c_alpha = d->c_alpha;
    c->vof.cfl = d->cfl_alpha;
    c->vof.a_in = d->alpha_inflow;
    c->vof.on = 1;
    TRY(launch_vof_fill(c, 0, -1));
    VofArgs m = vof_args(c, c->vof.alpha[0], c->vof.alpha[0], -1);
    hipLaunchKernelGGL(k_vof_monitors, dim3(c->red_blocks), dim3(256), 0, c->stream, m, c->geo);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(c->stream));
    return 0;
}

}  // extern "C"
