// Batched ensembles of lid-driven and heated cavities on the unit square
// (cfd2d/cases.py:60-70 and 207-223): many independent instances advanced in
// lockstep, instances of one side length n forming a class whose kernels take
// the instance from blockIdx.y.  The step is the one of the single-case path
// (k_bc_domain, k_dt_gamma, k_predictor, k_project, k_scalar_transport,
// k_fill_ghosts, k_record) restricted to closed cavities with no-slip walls,
// with the bulk in FP32: velocities, predictor, temperature, pressure and the
// V-cycle are FP32; the flexible CG keeps its iterate, residual, directions and
// every dot product in FP64, so each solve still ends on an FP64 true residual.
//
// Pressure solve: the learned flexible CG of lpcg.hip, batched.  Each class
// runs V-cycle legs down to its 32x32 level; all instances' 32x32 coarse
// right-hand sides go through the shared U-Net in one batched forward pass.
#pragma clang fp contract(off)
#include "cfd_kernels.h"

constexpr int ENS_MAXLEV = 12;
constexpr int ENS_BANDS = 64;     // row bands per instance of the reducing kernels
constexpr int ENS_NV = 4;         // partial-sum values per band
enum { ENS_P_A = 0, ENS_P_B = 1, ENS_P_C = 2, ENS_P_D = 3, ENS_NSLOT = 4 };

struct EnsLevel {
    int n, P;
    long S;                   // instance stride (elements)
    const float* diag;        // shared 5-point operator (levels >= 1)
    const float* cE;
    const float* cN;
    float *b, *x, *t;         // per instance
};

struct EnsState {
    double t, dt, dt_old, gamma, ab_a, ab_b, umax, vmax;
    double residual, ke, divmax;
    double bnorm, atol, rr, rho, rho_old, beta, alpha, relres;
    long long nstep;
    int has_old, k, first, done, pending, iters, info, reps;
};

// Step control shared by all classes (device memory).  A step captured as one
// graph runs a fixed number of solver iterations (finished instances return
// at once); when a solve is still unfinished after them, k_e_cap sets `stall`:
// every pre- and post-solve kernel then returns at once, so the next launch of
// the graph only continues the iterations, and the step ends in the launch
// whose k_e_cap finds every solve done.  The buffer parity and the history of
// the extrapolated start live here too, so a launch depends on no host value.
struct EnsCtl {
    int stall;               // a solve of the current step is unfinished
    int par;                 // parity of the step's hu / hv / T buffers
    int hist;                // the extrapolated start has its history
    int x0ext;               // extrapolated initial iterates in use
    long long iters;         // lockstep solver iterations of all completed steps
    int last_max;            // lockstep iterations of the last completed step
    int max_call;            // their largest value since the host cleared it
    int stalls;              // launches that ended with a solve unfinished
    int pad;
};

struct EnsClass {
    int n, P, B, K, nlev;
    long S;
    float h, ax;              // h = 1/n, ax = n^2
    double hd, axd;
    double tau;
    const double* nu;         // per instance
    const double* lid;
    const double* kappa;      // thermal diffusivity (0: no temperature)
    float *u, *v, *F, *G, *hu[2], *hv[2], *T[2], *p;
    double *x, *b, *r, *ro, *q, *d0, *d1;
    double* xp;               // previous solution (extrapolated initial iterate), or null
    EnsLevel lev[ENS_MAXLEV];
    int ncoarse;
    const float* cinv;
    float omega, wc;
    EnsState* st;
    double* part;             // [slot][B][ENS_BANDS][ENS_NV]
    int img0;                 // first image of the class in the network batch
    float* img;               // network input (all images)
    const float* y;           // network output
    float gain;               // sigma / D
    double one_a_one;         // 1^T A_K 1
    double rtol;
    int maxiter;
    int use_net;
    int* notdone;             // per class: instances still iterating
    EnsCtl* ctl;              // shared step control
};

namespace {

__device__ __forceinline__ double* ens_part(const EnsClass* c, int slot, int b) {
    return c->part + ((long)slot * c->B + b) * ENS_BANDS * ENS_NV;
}

// rows [i0, i1) of band `band` out of nb over rows lo..hi-1
__device__ __forceinline__ void ens_band(int lo, int hi, int band, int nb, int& i0, int& i1) {
    const long len = hi - lo;
    i0 = lo + (int)((len * band) / nb);
    i1 = lo + (int)((len * (band + 1)) / nb);
}

// the interior cells (i, j) of band blockIdx.x of an n x n instance, U cells
// per thread at a time: gather (loads and arithmetic, no stores) for all U
// cells before scatter stores any, so their loads are in flight together.
// Per thread the cells come in the order of the row loop (i, then j by
// blockDim.x) when n is a multiple of blockDim.x.
template <int U, typename G, typename S>
__device__ __forceinline__ void ens_cells_batch(int n, G gather, S scatter) {
    int i0, i1;
    ens_band(1, n + 1, blockIdx.x, ENS_BANDS, i0, i1);
    const int ncell = (i1 - i0) * n, bs = blockDim.x;
    using T = decltype(gather(1, 1));
    for (int q0 = threadIdx.x; q0 < ncell; q0 += U * bs) {
        T val[U];
        int ii[U], jj[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int q = q0 + u * bs < ncell ? q0 + u * bs : q0;
            ii[u] = i0 + q / n;
            jj[u] = 1 + q % n;
            val[u] = gather(ii[u], jj[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (q0 + u * bs < ncell) scatter(ii[u], jj[u], val[u]);
    }
}

template <int NV>
__device__ __forceinline__ void ens_partial(double (&v)[NV], double* slot, int band) {
    __shared__ double lds[NV * 16];
    block_sum<NV>(v, lds);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < NV; ++q) slot[band * ENS_NV + q] = v[q];
}

template <int NV>
__device__ __forceinline__ void ens_partial_max(double (&v)[NV], double* slot, int band) {
    __shared__ double lds[NV * 16];
    block_max<NV>(v, lds);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < NV; ++q) slot[band * ENS_NV + q] = v[q];
}

// totals of an instance's band partials, by the whole work-group (one band per
// thread, then a block sum; every thread of the work-group must call it)
template <int NV>
__device__ __forceinline__ void ens_total_wg(const double* slot, double (&v)[NV]) {
    __shared__ double lds[NV * 16];
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = 0.0;
    for (int k = threadIdx.x; k < ENS_BANDS; k += blockDim.x)
#pragma unroll
        for (int q = 0; q < NV; ++q) v[q] += slot[k * ENS_NV + q];
    block_sum<NV>(v, lds);
}

// the same totals by one thread (kernels with a thread per instance)
template <int NV>
__device__ __forceinline__ void ens_total(const double* slot, double (&v)[NV]) {
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = 0.0;
    for (int k = 0; k < ENS_BANDS; ++k)
#pragma unroll
        for (int q = 0; q < NV; ++q) v[q] += slot[k * ENS_NV + q];
}

// fine-level diagonal of the pinned Neumann cavity (pressure.py:53-125): the
// couplings to the neighbours inside the box, plus 1/h^2 at the pinned cell (1, 1)
__device__ __forceinline__ int diag_count(int i, int j, int n) {
    return (i > 1) + (i < n) + (j > 1) + (j < n) + (i == 1 && j == 1);
}

__device__ __forceinline__ float lev_diag(const EnsClass* c, int l, const EnsLevel& L, int i,
                                          int j) {
    if (l == 0) return c->ax * (float)diag_count(i, j, c->n);
    return L.diag[ix(i, j, L.P)];
}

// (A v)(i, j) on level l of a class; v(i, j) is 0 outside the box
template <typename Fv>
__device__ __forceinline__ float lev_apply(const EnsClass* c, int l, const EnsLevel& L, int i,
                                           int j, float d, Fv val) {
    if (l == 0)
        return d * val(i, j) - c->ax * (val(i + 1, j) + val(i - 1, j)) -
               c->ax * (val(i, j + 1) + val(i, j - 1));
    const int P = L.P;
    return d * val(i, j) - L.cE[ix(i, j, P)] * val(i + 1, j) - L.cE[ix(i - 1, j, P)] * val(i - 1, j) -
           L.cN[ix(i, j, P)] * val(i, j + 1) - L.cN[ix(i, j - 1, P)] * val(i, j - 1);
}

// momentum tendencies at one face (kernels.py:84-155), FP32
__device__ __forceinline__ float mom_u32(const float* u, const float* v, int P, int i, int j,
                                         float nu, float gamma, float idx2, float idx) {
    const float uc = u[ix(i, j, P)], ue = u[ix(i + 1, j, P)], uw = u[ix(i - 1, j, P)];
    const float un = u[ix(i, j + 1, P)], us = u[ix(i, j - 1, P)];
    const float diff = nu * ((ue - 2.0f * uc + uw) * idx2 + (un - 2.0f * uc + us) * idx2);
    const float a_e = 0.5f * (uc + ue), a_w = 0.5f * (uw + uc);
    const float du2dx = idx * (a_e * a_e - a_w * a_w)
        + gamma * idx * (fabsf(a_e) * 0.5f * (uc - ue) - fabsf(a_w) * 0.5f * (uw - uc));
    const float v_n = 0.5f * (v[ix(i, j, P)] + v[ix(i + 1, j, P)]);
    const float v_s = 0.5f * (v[ix(i, j - 1, P)] + v[ix(i + 1, j - 1, P)]);
    const float duvdy = idx * (v_n * 0.5f * (uc + un) - v_s * 0.5f * (us + uc))
        + gamma * idx * (fabsf(v_n) * 0.5f * (uc - un) - fabsf(v_s) * 0.5f * (us - uc));
    return diff - du2dx - duvdy;
}

__device__ __forceinline__ float mom_v32(const float* u, const float* v, int P, int i, int j,
                                         float nu, float gamma, float idx2, float idx) {
    const float vc = v[ix(i, j, P)], ve = v[ix(i + 1, j, P)], vw = v[ix(i - 1, j, P)];
    const float vn = v[ix(i, j + 1, P)], vs = v[ix(i, j - 1, P)];
    const float diff = nu * ((ve - 2.0f * vc + vw) * idx2 + (vn - 2.0f * vc + vs) * idx2);
    const float b_n = 0.5f * (vc + vn), b_s = 0.5f * (vs + vc);
    const float dv2dy = idx * (b_n * b_n - b_s * b_s)
        + gamma * idx * (fabsf(b_n) * 0.5f * (vc - vn) - fabsf(b_s) * 0.5f * (vs - vc));
    const float u_e = 0.5f * (u[ix(i, j, P)] + u[ix(i, j + 1, P)]);
    const float u_w = 0.5f * (u[ix(i - 1, j, P)] + u[ix(i - 1, j + 1, P)]);
    const float duvdx = idx * (u_e * 0.5f * (vc + ve) - u_w * 0.5f * (vw + vc))
        + gamma * idx * (fabsf(u_e) * 0.5f * (vc - ve) - fabsf(u_w) * 0.5f * (vw - vc));
    return diff - duvdx - dv2dy;
}

__device__ __forceinline__ float vanleer(float r) { return r <= 0.0f ? 0.0f : 2.0f * r / (1.0f + r); }

__device__ __forceinline__ float tvd_face(float phu, float phc, float phd, bool has_u) {
    if (!has_u) return phc;
    const float den = phd - phc;
    if (fabsf(den) < 1e-30f) return phc;
    const float r = (phc - phu) / den;
    return phc + 0.5f * vanleer(r) * den;
}

}  // namespace

// ---- explicit part of the step -------------------------------------------
// cavity walls (boundary.py:143-223 for four no-slip walls, lid on top), one
// work-group per instance, in the reference's statement order
__global__ void __launch_bounds__(256) k_e_bc(const EnsClass* __restrict__ c) {
    if (c->ctl->stall) return;
    const int b = blockIdx.x, n = c->n, P = c->P;
    float* u = c->u + b * c->S;
    float* v = c->v + b * c->S;
    const float lid = (float)c->lid[b];
    for (int j = threadIdx.x; j <= n; j += blockDim.x) {
        if (j >= 1) {
            u[ix(0, j, P)] = 0.0f;
            u[ix(n, j, P)] = 0.0f;
        }
        v[ix(0, j, P)] = 2.0f * 0.0f - v[ix(1, j, P)];
        v[ix(n + 1, j, P)] = 2.0f * 0.0f - v[ix(n, j, P)];
    }
    __syncthreads();
    for (int i = threadIdx.x; i <= n; i += blockDim.x) {
        if (i >= 1) {
            v[ix(i, 0, P)] = 0.0f;
            v[ix(i, n, P)] = 0.0f;
        }
        u[ix(i, 0, P)] = 2.0f * 0.0f - u[ix(i, 1, P)];
        u[ix(i, n + 1, P)] = 2.0f * lid - u[ix(i, n, P)];
    }
}

// max |u|, max |v| over the whole arrays (ghosts included), per band
__global__ void __launch_bounds__(256) k_e_maxuv(const EnsClass* __restrict__ c) {
    if (c->ctl->stall) return;
    const int b = blockIdx.y, n = c->n, P = c->P;
    const float* u = c->u + b * c->S;
    const float* v = c->v + b * c->S;
    int i0, i1;
    ens_band(0, n + 2, blockIdx.x, ENS_BANDS, i0, i1);
    double m[2] = {0.0, 0.0};
    for (int i = i0; i < i1; ++i)
        for (int j = threadIdx.x; j <= n + 1; j += blockDim.x) {
            if (i <= n) m[0] = nanmax(m[0], (double)fabsf(u[ix(i, j, P)]));
            if (j <= n) m[1] = nanmax(m[1], (double)fabsf(v[ix(i, j, P)]));
        }
    ens_partial_max<2>(m, ens_part(c, ENS_P_A, b), blockIdx.x);
}

// dt, gamma and the AB2 weights of every instance (solver.py:209-253)
__global__ void __launch_bounds__(256) k_e_dt(const EnsClass* __restrict__ c) {
    if (c->ctl->stall) return;
    for (int b = threadIdx.x; b < c->B; b += blockDim.x) {
        const double* pa = ens_part(c, ENS_P_A, b);
        double umax_raw = 0.0, vmax_raw = 0.0;
        for (int k = 0; k < ENS_BANDS; ++k) {
            umax_raw = nanmax(umax_raw, pa[k * ENS_NV]);
            vmax_raw = nanmax(vmax_raw, pa[k * ENS_NV + 1]);
        }
        EnsState& s = c->st[b];
        const double h = c->hd, nu = c->nu[b];
        const double umax = fmax(umax_raw, 1e-12), vmax = fmax(vmax_raw, 1e-12);
        const double dt_visc = 0.5 * (0.5 / (nu * (1.0 / (h * h) + 1.0 / (h * h))));
        const double dt_conv = 0.8 * fmin(h / umax, h / vmax);
        double dt = c->tau * fmin(dt_visc, dt_conv);
        const double kap = c->kappa[b];
        if (kap > 0.0) dt = fmin(dt, c->tau * (0.5 / (kap * (1.0 / (h * h) + 1.0 / (h * h)))));
        const double re_cell = fmax(umax_raw * h, vmax_raw * h) / nu;
        s.gamma = fmin(1.0, fmax(0.05, 0.015 * re_cell));
        s.umax = umax_raw;
        s.vmax = vmax_raw;
        s.dt = dt;
        if (s.has_old) {
            const double r = dt / s.dt_old;
            s.ab_a = 1.0 + 0.5 * r;
            s.ab_b = -0.5 * r;
        } else {
            s.ab_a = 1.0;
            s.ab_b = 0.0;
        }
    }
}

// fused predictor: tendencies + Boussinesq buoyancy (gravity (0, -1),
// T_ref 0) + AB2 + F/G; writes hu/hv of parity par
__global__ void k_e_pred(const EnsClass* __restrict__ c) {
    if (c->ctl->stall) return;
    const int par = c->ctl->par;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    const int b = blockIdx.z, n = c->n, P = c->P;
    if (i > n + 1 || j > n + 1) return;
    const long o = b * c->S;
    const float* u = c->u + o;
    const float* v = c->v + o;
    const EnsState& s = c->st[b];
    const float dt = (float)s.dt, gamma = (float)s.gamma;
    const float A = (float)s.ab_a, Bw = (float)s.ab_b;
    const bool ab2 = s.has_old;
    const float nu = (float)c->nu[b];
    const float idx = (float)c->n, idx2 = c->ax;
    const float* T = c->kappa[b] > 0.0 ? c->T[par] + o : nullptr;
    if (i <= n) {
        const bool uf = i >= 1 && i <= n - 1 && j >= 1 && j <= n;
        float h = 0.0f;
        if (uf) h = mom_u32(u, v, P, i, j, nu, gamma, idx2, idx);
        const float uval = u[ix(i, j, P)];
        float F = uval;
        if (uf) {
            const float H = ab2 ? A * h + Bw * c->hu[par ^ 1][o + ix(i, j, P)] : h;
            F = uval + dt * H;
        }
        c->hu[par][o + ix(i, j, P)] = h;
        c->F[o + ix(i, j, P)] = F;
    }
    if (j <= n) {
        const bool vf = i >= 1 && i <= n && j >= 1 && j <= n - 1;
        float h = 0.0f;
        if (vf) h = mom_v32(u, v, P, i, j, nu, gamma, idx2, idx);
        if (vf && T) {
            const float Tf = 0.5f * (T[ix(i, j, P)] + T[ix(i, j + 1, P)]) - 0.0f;
            h += -(-1.0f) * Tf;
        }
        const float vval = v[ix(i, j, P)];
        float G = vval;
        if (vf) {
            const float H = ab2 ? A * h + Bw * c->hv[par ^ 1][o + ix(i, j, P)] : h;
            G = vval + dt * H;
        }
        c->hv[par][o + ix(i, j, P)] = h;
        c->G[o + ix(i, j, P)] = G;
    }
}

// right-hand side div(F, G) / dt (FP64) and its partial sums
__global__ void __launch_bounds__(256) k_e_rhs(const EnsClass* __restrict__ c) {
    if (c->ctl->stall) return;
    const int b = blockIdx.y, n = c->n, P = c->P;
    const long o = b * c->S;
    const double dt = c->st[b].dt, h = c->hd;
    int i0, i1;
    ens_band(1, n + 1, blockIdx.x, ENS_BANDS, i0, i1);
    double s[1] = {0.0};
    for (int i = i0; i < i1; ++i)
        for (int j = 1 + threadIdx.x; j <= n; j += blockDim.x) {
            const double div = ((double)c->F[o + ix(i, j, P)] - (double)c->F[o + ix(i - 1, j, P)]) / h +
                               ((double)c->G[o + ix(i, j, P)] - (double)c->G[o + ix(i, j - 1, P)]) / h;
            const double rhs = div / dt;
            c->b[o + ix(i, j, P)] = rhs;
            s[0] += rhs;
        }
    ens_partial<1>(s, ens_part(c, ENS_P_A, b), blockIdx.x);
}

// initial iterate of the instance's solve: x <- x + theta (x - x_prev) with
// theta = dt / dt_old, x_prev <- the old x (hist = 0: the copy only)
__global__ void __launch_bounds__(256) k_e_x0(const EnsClass* __restrict__ c) {
    if (c->ctl->stall) return;
    const int hist = c->ctl->hist;
    const int b = blockIdx.y, n = c->n, P = c->P;
    const EnsState& st = c->st[b];
    const double th = hist && st.dt_old > 0.0 ? st.dt / st.dt_old : 0.0;
    const long o = b * c->S;
    int i0, i1;
    ens_band(1, n + 1, blockIdx.x, ENS_BANDS, i0, i1);
    for (int i = i0; i < i1; ++i)
        for (int j = 1 + threadIdx.x; j <= n; j += blockDim.x) {
            const long k = o + ix(i, j, P);
            const double xn = c->x[k];
            if (hist) c->x[k] = xn + th * (xn - c->xp[k]);
            c->xp[k] = xn;
        }
}

// b = -(rhs - mean) (compatibility), r = b - A x, first V-cycle input;
// partial b.b and r.r
__global__ void __launch_bounds__(256) k_e_res(const EnsClass* __restrict__ c) {
    if (c->ctl->stall) return;
    const int b = blockIdx.y, n = c->n, P = c->P;
    const long o = b * c->S;
    double tot[1];
    ens_total_wg<1>(ens_part(c, ENS_P_A, b), tot);
    const double mean = tot[0] / ((double)n * n);
    const EnsLevel& L0 = c->lev[0];
    int i0, i1;
    ens_band(1, n + 1, blockIdx.x, ENS_BANDS, i0, i1);
    double s[2] = {0.0, 0.0};
    const double* x = c->x + o;
    for (int i = i0; i < i1; ++i)
        for (int j = 1 + threadIdx.x; j <= n; j += blockDim.x) {
            const long k = o + ix(i, j, P);
            const double bv = -(c->b[k] - mean);
            c->b[k] = bv;
            const double d = c->axd * diag_count(i, j, n);
            const double Ax = d * x[ix(i, j, P)] - c->axd * (x[ix(i + 1, j, P)] + x[ix(i - 1, j, P)]) -
                              c->axd * (x[ix(i, j + 1, P)] + x[ix(i, j - 1, P)]);
            const double rv = bv - Ax;
            c->r[k] = rv;
            s[0] += bv * bv;
            s[1] += rv * rv;
            const float rf = (float)rv;
            L0.b[b * L0.S + ix(i, j, P)] = rf;
            L0.x[b * L0.S + ix(i, j, P)] = c->omega * rf / (c->ax * (float)diag_count(i, j, n));
        }
    ens_partial<2>(s, ens_part(c, ENS_P_B, b), blockIdx.x);
}

// per instance: the solver's initial state
__global__ void __launch_bounds__(256) k_e_init(const EnsClass* __restrict__ c) {
    if (c->ctl->stall) return;
    int cnt = 0;
    for (int b = threadIdx.x; b < c->B; b += blockDim.x) {
        double s[2];
        ens_total<2>(ens_part(c, ENS_P_B, b), s);
        EnsState& st = c->st[b];
        st.bnorm = sqrt(s[0]);
        st.atol = c->rtol * st.bnorm;
        st.rr = s[1];
        st.rho_old = 1.0;
        st.k = 0;
        st.first = 1;
        st.pending = 0;
        st.iters = 0;
        st.info = 0;
        st.reps = 0;
        st.done = (st.bnorm == 0.0 || sqrt(s[1]) < st.atol) ? 1 : 0;
        cnt += st.done ? 0 : 1;
    }
    __shared__ int s_cnt;
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    atomicAdd(&s_cnt, cnt);
    __syncthreads();
    if (threadIdx.x == 0) *c->notdone = s_cnt;
}

// ---- batched V-cycle and flexible CG -------------------------------------
// one cell per thread of level l: work-group = a tile of blockDim.y rows x
// blockDim.x columns (the shape the autotuner picked for the class and level;
// every shape gives the same values), blockIdx.x = tile (row-major over the
// level), y = instance
#define ENS_CELL(L)                                                           \
    const int cj_ = ((L).n + (int)blockDim.x - 1) / (int)blockDim.x;          \
    const int ti_ = blockIdx.x / cj_;                                         \
    const int i = ti_ * (int)blockDim.y + (int)threadIdx.y + 1;               \
    const int j = (blockIdx.x - ti_ * cj_) * (int)blockDim.x + (int)threadIdx.x + 1; \
    if (i > (L).n || j > (L).n) return

__global__ void __launch_bounds__(256) k_e_smooth(const EnsClass* __restrict__ c, int l,
                                                  int from_t) {
    const int b = blockIdx.y;
    if (c->st[b].done) return;
    const EnsLevel& L = c->lev[l];
    ENS_CELL(L);
    const long o = b * L.S;
    const float* src = (from_t ? L.t : L.x) + o;
    float* dst = (from_t ? L.x : L.t) + o;
    const float d = lev_diag(c, l, L, i, j);
    if (d == 0.0f) return;
    auto sv = [&](int p, int q) -> float { return src[ix(p, q, L.P)]; };
    dst[ix(i, j, L.P)] = src[ix(i, j, L.P)] + c->omega * (L.b[o + ix(i, j, L.P)] -
                                                          lev_apply(c, l, L, i, j, d, sv)) / d;
}

__global__ void __launch_bounds__(256) k_e_restrict(const EnsClass* __restrict__ c, int l) {
    const int b = blockIdx.y;
    if (c->st[b].done) return;
    const EnsLevel& L = c->lev[l];
    const EnsLevel& C = c->lev[l + 1];
    const int cj_ = (C.n + (int)blockDim.x - 1) / (int)blockDim.x;
    const int ti_ = blockIdx.x / cj_;
    const int ci = ti_ * (int)blockDim.y + (int)threadIdx.y + 1;
    const int cjj = (blockIdx.x - ti_ * cj_) * (int)blockDim.x + (int)threadIdx.x + 1;
    if (ci > C.n || cjj > C.n) return;
    const long o = b * L.S, oc = b * C.S;
    float bc = 0.0f;
    for (int di = 0; di < 2; ++di)
        for (int dj = 0; dj < 2; ++dj) {
            const int i = 2 * ci - 1 + di, j = 2 * cjj - 1 + dj;
            const float d = lev_diag(c, l, L, i, j);
            if (d == 0.0f) continue;
            auto tv = [&](int p, int q) -> float { return L.t[o + ix(p, q, L.P)]; };
            bc += L.b[o + ix(i, j, L.P)] - lev_apply(c, l, L, i, j, d, tv);
        }
    const long cc = ix(ci, cjj, C.P);
    C.b[oc + cc] = bc;
    const float dc = C.diag[cc];
    C.x[oc + cc] = dc != 0.0f ? c->omega * bc / dc : 0.0f;
}

__global__ void __launch_bounds__(256) k_e_dense(const EnsClass* __restrict__ c) {
    const int b = blockIdx.x;
    if (c->st[b].done) return;
    const EnsLevel& C = c->lev[c->nlev - 1];
    const int n = c->ncoarse;
    const long o = b * C.S;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        float acc = 0.0f;
        for (int q = 0; q < n; ++q)
            acc += c->cinv[(long)k * n + q] * C.b[o + ix(q / C.n + 1, q % C.n + 1, C.P)];
        C.t[o + ix(k / C.n + 1, k % C.n + 1, C.P)] = acc;
    }
}

__global__ void __launch_bounds__(256) k_e_post(const EnsClass* __restrict__ c, int l) {
    const int b = blockIdx.y;
    if (c->st[b].done) return;
    const EnsLevel& L = c->lev[l];
    const EnsLevel& C = c->lev[l + 1];
    ENS_CELL(L);
    const long o = b * L.S, oc = b * C.S;
    const float wc = c->wc;
    auto corr = [&](int p, int q) -> float {
        if (p < 1 || p > L.n || q < 1 || q > L.n) return 0.0f;
        return L.t[o + ix(p, q, L.P)] + wc * C.t[oc + ix((p + 1) >> 1, (q + 1) >> 1, C.P)];
    };
    const float d = lev_diag(c, l, L, i, j);
    if (d == 0.0f) return;
    const float y = corr(i, j);
    L.x[o + ix(i, j, L.P)] = y + c->omega * (L.b[o + ix(i, j, L.P)] - lev_apply(c, l, L, i, j, d, corr)) / d;
}

// level K sums for the constant-vector correction (partials per band)
__global__ void __launch_bounds__(256) k_e_cmode(const EnsClass* __restrict__ c) {
    const int b = blockIdx.y;
    if (c->st[b].done) return;
    const int l = c->K;
    const EnsLevel& L = c->lev[l];
    const long o = b * L.S;
    int i0, i1;
    ens_band(1, L.n + 1, blockIdx.x, ENS_BANDS, i0, i1);
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    auto one = [&](int p, int q) -> float { return (p >= 1 && p <= L.n && q >= 1 && q <= L.n) ? 1.0f : 0.0f; };
    for (int i = i0; i < i1; ++i)
        for (int j = 1 + threadIdx.x; j <= L.n; j += blockDim.x) {
            const float d = lev_diag(c, l, L, i, j);
            if (d == 0.0f) continue;
            const double a1 = lev_apply(c, l, L, i, j, d, one);
            const double bv = L.b[o + ix(i, j, L.P)];
            s[0] += bv;
            s[1] += bv * bv;
            s[2] += bv * a1;
            s[3] += a1 * a1;
        }
    ens_partial<4>(s, ens_part(c, ENS_P_C, b), blockIdx.x);
}

__device__ __forceinline__ void ens_cstats(const EnsClass* c, int b, double& cm, double& s_in) {
    double s[4];
    ens_total_wg<4>(ens_part(c, ENS_P_C, b), s);
    const double nf = (double)c->lev[c->K].n * c->lev[c->K].n;
    cm = c->one_a_one > 0.0 ? s[0] / c->one_a_one : 0.0;
    const double bb = s[1] - 2.0 * cm * s[2] + cm * cm * s[3];
    s_in = sqrt(fmax(bb, 0.0) / nf);
}

// network input image of instance b (32 x 32 at rows 32 (img0 + b) of the batch)
__global__ void __launch_bounds__(256) k_e_net_in(const EnsClass* __restrict__ c) {
    const int b = blockIdx.y;
    const int l = c->K;
    const EnsLevel& L = c->lev[l];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;   // 0 .. 1023
    const bool done = c->st[b].done;
    double cm = 0.0, s = 0.0;
    if (!done) ens_cstats(c, b, cm, s);   // the whole work-group, before any thread leaves
    if (k >= L.n * L.n) return;
    const int i = k / L.n + 1, j = k % L.n + 1;
    float val = 0.0f;
    if (!done) {
        const float d = lev_diag(c, l, L, i, j);
        auto one = [&](int p, int q) -> float { return (p >= 1 && p <= L.n && q >= 1 && q <= L.n) ? 1.0f : 0.0f; };
        if (d != 0.0f && s > 0.0)
            val = (float)(((double)L.b[b * L.S + ix(i, j, L.P)] - cm * (double)lev_apply(c, l, L, i, j, d, one)) / s);
    }
    c->img[(long)(c->img0 + b) * L.n * L.n + k] = val;
}

// level-K correction t = c + gain s y, then (separate launches) two sweeps
__global__ void __launch_bounds__(256) k_e_net_out(const EnsClass* __restrict__ c) {
    const int b = blockIdx.y;
    if (c->st[b].done) return;
    const int l = c->K;
    const EnsLevel& L = c->lev[l];
    double cm, s;
    ens_cstats(c, b, cm, s);   // the whole work-group, before any thread leaves
    ENS_CELL(L);
    double t = cm;
    if (c->use_net)
        t += (double)c->gain * s * (double)c->y[(long)(c->img0 + b) * L.n * L.n + (i - 1) * L.n + (j - 1)];
    L.t[b * L.S + ix(i, j, L.P)] = (float)t;
}

// last sweep on the fine level with the partial dot products r.z, r_old.z
template <int U>
__global__ void __launch_bounds__(256) k_e_smooth_dots(const EnsClass* __restrict__ c) {
    const int b = blockIdx.y;
    const EnsState& st = c->st[b];
    if (st.done) return;
    const EnsLevel& L = c->lev[0];
    const int n = c->n, P = c->P;
    const long o = b * L.S;
    const bool first = st.first;
    double s[2] = {0.0, 0.0};
    struct Z {
        float z;
        double r, ro;
    };
    ens_cells_batch<U>(
        n,
        [&](int i, int j) {
            const float d = lev_diag(c, 0, L, i, j);
            auto xv = [&](int p, int q) -> float { return L.x[o + ix(p, q, P)]; };
            Z w;
            w.z = L.x[o + ix(i, j, P)] +
                  c->omega * (L.b[o + ix(i, j, P)] - lev_apply(c, 0, L, i, j, d, xv)) / d;
            const long k = b * c->S + ix(i, j, P);
            w.r = c->r[k];
            w.ro = c->ro[k];
            return w;
        },
        [&](int i, int j, const Z& w) {
            L.t[o + ix(i, j, P)] = w.z;
            s[0] += w.r * (double)w.z;
            if (!first) s[1] += w.ro * (double)w.z;
        });
    ens_partial<2>(s, ens_part(c, ENS_P_A, b), blockIdx.x);
}

// d = z + beta d_old, q = A d, partial d.q
template <int U>
__global__ void __launch_bounds__(256) k_e_dir(const EnsClass* __restrict__ c) {
    const int b = blockIdx.y;
    EnsState& st = c->st[b];
    if (st.done) return;
    double dots[2];
    ens_total_wg<2>(ens_part(c, ENS_P_A, b), dots);
    const bool first = st.first;
    const double beta = first ? 0.0 : (dots[0] - dots[1]) / st.rho_old;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st.rho = dots[0];
        st.beta = beta;
    }
    const int n = c->n, P = c->P;
    const long o = b * c->S;
    const float* z = c->lev[0].t + b * c->lev[0].S;
    double* dnew = ((st.k & 1) ? c->d1 : c->d0) + o;
    const double* dold = ((st.k & 1) ? c->d0 : c->d1) + o;
    double pq[1] = {0.0};
    struct D {
        double dc, qv;
    };
    // z and both direction buffers are 0 on the ghost ring (zeroed at
    // creation, only interior cells written), so d there is 0 as well; the
    // two variants keep the loads free of branches
    auto sweep = [&](auto first_c) {
        constexpr bool FIRST = decltype(first_c)::value;
        ens_cells_batch<U>(
            n,
            [&](int i, int j) {
                auto dirv = [&](int p, int q) -> double {
                    const long k = ix(p, q, P);
                    return FIRST ? (double)z[k] : (double)z[k] + beta * dold[k];
                };
                D w;
                w.dc = dirv(i, j);
                w.qv = c->axd * diag_count(i, j, n) * w.dc - c->axd * (dirv(i + 1, j) + dirv(i - 1, j)) -
                       c->axd * (dirv(i, j + 1) + dirv(i, j - 1));
                return w;
            },
            [&](int i, int j, const D& w) {
                dnew[ix(i, j, P)] = w.dc;
                c->q[o + ix(i, j, P)] = w.qv;
                pq[0] += w.dc * w.qv;
            });
    };
    if (first) sweep(std::true_type{});
    else sweep(std::false_type{});
    ens_partial<1>(pq, ens_part(c, ENS_P_B, b), blockIdx.x);
}

// alpha = rho / d.q; x += alpha d, r_old = r, r -= alpha q; next V-cycle input;
// partial r.r
template <int U>
__global__ void __launch_bounds__(256) k_e_update(const EnsClass* __restrict__ c) {
    const int b = blockIdx.y;
    EnsState& st = c->st[b];
    if (st.done) return;
    double pq[1];
    ens_total_wg<1>(ens_part(c, ENS_P_B, b), pq);
    const double alpha = st.rho / pq[0];
    if (!(pq[0] > 0.0) || !(fabs(alpha) < 1e300)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            st.done = 1;
            st.info = -1;
        }
        return;
    }
    const int n = c->n, P = c->P;
    const long o = b * c->S;
    const EnsLevel& L0 = c->lev[0];
    const double* dnew = ((st.k & 1) ? c->d1 : c->d0) + o;
    double s[1] = {0.0};
    struct Up {
        double x, r, rn;
    };
    ens_cells_batch<U>(
        n,
        [&](int i, int j) {
            const long k = o + ix(i, j, P);
            Up w;
            w.x = c->x[k] + alpha * dnew[ix(i, j, P)];
            w.r = c->r[k];
            w.rn = w.r - alpha * c->q[k];
            return w;
        },
        [&](int i, int j, const Up& w) {
            const long k = o + ix(i, j, P);
            c->x[k] = w.x;
            c->ro[k] = w.r;
            c->r[k] = w.rn;
            s[0] += w.rn * w.rn;
            const float rf = (float)w.rn;
            L0.b[b * L0.S + ix(i, j, P)] = rf;
            L0.x[b * L0.S + ix(i, j, P)] = c->omega * rf / (c->ax * (float)diag_count(i, j, n));
        });
    ens_partial<1>(s, ens_part(c, ENS_P_C, b), blockIdx.x);
}

// recursive-residual test of the iteration (every band; band 0 records), and
// the true residual r = b - A x once it passes or at maxiter; partial r.r
template <int U>
__global__ void __launch_bounds__(256) k_e_true(const EnsClass* __restrict__ c) {
    const int b = blockIdx.y;
    EnsState& st = c->st[b];
    if (st.done) return;
    double rr[1];
    ens_total_wg<1>(ens_part(c, ENS_P_C, b), rr);
    const int k = st.k + 1;   // st.k advances in k_e_true_fin
    const bool nan = !(rr[0] == rr[0]);
    const bool pending = !nan && (sqrt(rr[0]) < st.atol || k >= c->maxiter);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st.rr = rr[0];
        st.rho_old = st.rho;
        st.iters = k;
        st.first = 0;
        st.pending = pending ? 1 : 0;
        if (nan) {
            st.done = 1;
            st.info = -1;
        }
    }
    if (!pending) return;
    const int n = c->n, P = c->P;
    const long o = b * c->S;
    const EnsLevel& L0 = c->lev[0];
    const double* x = c->x + o;
    double s[1] = {0.0};
    ens_cells_batch<U>(
        n,
        [&](int i, int j) {
            const double d = c->axd * diag_count(i, j, n);
            const double Ax = d * x[ix(i, j, P)] - c->axd * (x[ix(i + 1, j, P)] + x[ix(i - 1, j, P)]) -
                              c->axd * (x[ix(i, j + 1, P)] + x[ix(i, j - 1, P)]);
            return c->b[o + ix(i, j, P)] - Ax;
        },
        [&](int i, int j, double rt) {
            c->r[o + ix(i, j, P)] = rt;
            s[0] += rt * rt;
            const float rf = (float)rt;
            L0.b[b * L0.S + ix(i, j, P)] = rf;
            L0.x[b * L0.S + ix(i, j, P)] = c->omega * rf / (c->ax * (float)diag_count(i, j, n));
        });
    ens_partial<1>(s, ens_part(c, ENS_P_D, b), blockIdx.x);
}

// per instance: advance the iteration count; after a true residual, finish or
// restart; count the instances still iterating
__global__ void __launch_bounds__(256) k_e_true_fin(const EnsClass* __restrict__ c) {
    int cnt = 0;
    for (int b = threadIdx.x; b < c->B; b += blockDim.x) {
        EnsState& st = c->st[b];
        if (st.done) continue;
        st.k += 1;
        if (st.pending) {
            double rr[1];
            ens_total<1>(ens_part(c, ENS_P_D, b), rr);
            st.rr = rr[0];
            st.reps += 1;
            st.pending = 0;
            if (sqrt(rr[0]) < st.atol) {
                st.done = 1;
                st.info = 0;
            } else if (!(rr[0] == rr[0])) {
                st.done = 1;
                st.info = -1;
            } else if (st.k >= c->maxiter) {
                st.done = 1;
                st.info = c->maxiter;
            } else {
                st.first = 1;
            }
        }
        cnt += st.done ? 0 : 1;
    }
    __shared__ int s_cnt;
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    atomicAdd(&s_cnt, cnt);
    __syncthreads();
    if (threadIdx.x == 0) *c->notdone = s_cnt;
}

// ---- end of the projection ------------------------------------------------
__global__ void __launch_bounds__(256) k_e_xsum(const EnsClass* __restrict__ c) {
    if (c->ctl->stall) return;
    const int b = blockIdx.y, n = c->n, P = c->P;
    const long o = b * c->S;
    int i0, i1;
    ens_band(1, n + 1, blockIdx.x, ENS_BANDS, i0, i1);
    double s[1] = {0.0};
    for (int i = i0; i < i1; ++i)
        for (int j = 1 + threadIdx.x; j <= n; j += blockDim.x) s[0] += c->x[o + ix(i, j, P)];
    ens_partial<1>(s, ens_part(c, ENS_P_A, b), blockIdx.x);
}

// pressure = x minus its mean (the pure-Neumann component), FP32
__global__ void __launch_bounds__(256) k_e_pwrite(const EnsClass* __restrict__ c) {
    if (c->ctl->stall) return;
    const int b = blockIdx.y, n = c->n, P = c->P;
    const long o = b * c->S;
    double tot[1];
    ens_total_wg<1>(ens_part(c, ENS_P_A, b), tot);
    const double mean = tot[0] / ((double)n * n);
    int i0, i1;
    ens_band(1, n + 1, blockIdx.x, ENS_BANDS, i0, i1);
    const bool zero = c->st[b].bnorm == 0.0;   // SciPy returns b itself (all zeros)
    for (int i = i0; i < i1; ++i)
        for (int j = 1 + threadIdx.x; j <= n; j += blockDim.x) {
            if (zero) c->x[o + ix(i, j, P)] = 0.0;
            c->p[o + ix(i, j, P)] = zero ? 0.0f : (float)(c->x[o + ix(i, j, P)] - mean);
        }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        EnsState& st = c->st[b];
        st.relres = st.bnorm > 0.0 ? sqrt(st.rr) / st.bnorm : 0.0;
    }
}

// u = F - dt grad p on interior faces (F elsewhere), partial max |u_new - u_old|
__global__ void __launch_bounds__(256) k_e_correct(const EnsClass* __restrict__ c) {
    if (c->ctl->stall) return;
    const int b = blockIdx.y, n = c->n, P = c->P;
    const long o = b * c->S;
    const float dt = (float)c->st[b].dt, h = c->h;
    const float* p = c->p + o;
    float* u = c->u + o;
    float* v = c->v + o;
    int i0, i1;
    ens_band(1, n + 1, blockIdx.x, ENS_BANDS, i0, i1);
    double res[1] = {0.0};
    for (int i = i0; i < i1; ++i) {
        for (int j = 1 + threadIdx.x; j <= n; j += blockDim.x) {
            const long k = ix(i, j, P);
            const float F = c->F[o + k];
            float un = F;
            if (i < n) {
                un = F - dt * (p[ix(i + 1, j, P)] - p[k]) / h;
                res[0] = nanmax(res[0], (double)fabsf(un - u[k]));
            }
            u[k] = un;
        }
        for (int j = threadIdx.x; j <= n; j += blockDim.x) {
            const long k = ix(i, j, P);
            const float G = c->G[o + k];
            float vn = G;
            if (j >= 1 && j < n) {
                vn = G - dt * (p[ix(i, j + 1, P)] - p[k]) / h;
                res[0] = nanmax(res[0], (double)fabsf(vn - v[k]));
            }
            v[k] = vn;
        }
    }
    if (blockIdx.x == 0)
        for (int j = 1 + threadIdx.x; j <= n; j += blockDim.x) u[ix(0, j, P)] = c->F[o + ix(0, j, P)];
    ens_partial_max<1>(res, ens_part(c, ENS_P_B, b), blockIdx.x);
}

// temperature: TVD (van Leer) transport and Euler update on every cell, with
// fixed values on the left (0.5) and right (-0.5) walls and zero gradient on
// the bottom and top (cases.py:207-223; kernels.py:254-391 for these codes)
__global__ void k_e_temp(const EnsClass* __restrict__ c) {
    if (c->ctl->stall) return;
    const int par = c->ctl->par;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    const int b = blockIdx.z, n = c->n, P = c->P;
    const float D = (float)c->kappa[b];
    if (D <= 0.0f || i > n + 1 || j > n + 1) return;
    const long o = b * c->S;
    const float* phi = c->T[par] + o;
    float* out = c->T[par ^ 1] + o;
    const float* u = c->u + o;
    const float* v = c->v + o;
    const float h = c->h;
    if (!(i >= 1 && i <= n && j >= 1 && j <= n)) {
        out[ix(i, j, P)] = phi[ix(i, j, P)];
        return;
    }
    // signed contribution of x-face f to cell `cell`
    auto xterm = [&](int f, int cell) -> float {
        if (f == 0 || f == n) {
            const float sgn = f == 0 ? 1.0f : -1.0f;
            const float val = f == 0 ? 0.5f : -0.5f;
            const float un = sgn * u[ix(f, j, P)];
            const float php = phi[ix(cell, j, P)];
            const float phb = un > 0.0f ? val : php;
            float flux_in = un * phb;
            flux_in += D * (val - php) / (0.5f * h);
            return flux_in / h;
        }
        const int L = f, R = f + 1;
        const float uf = u[ix(f, j, P)];
        int C, Dn, Ui;
        if (uf >= 0.0f) { C = L; Dn = R; Ui = L - 1; if (Ui < 1) Ui = -1; }
        else { C = R; Dn = L; Ui = R + 1; if (Ui > n) Ui = -1; }
        const bool has_u = Ui >= 1;
        const float phu = has_u ? phi[ix(Ui, j, P)] : 0.0f;
        const float phf = tvd_face(phu, phi[ix(C, j, P)], phi[ix(Dn, j, P)], has_u);
        const float df = D + D > 0.0f ? 2.0f * D * D / (D + D) : 0.0f;
        const float Fx = uf * phf - df * (phi[ix(R, j, P)] - phi[ix(L, j, P)]) / h;
        return cell == L ? -(Fx / h) : Fx / h;
    };
    auto yterm = [&](int f, int cell) -> float {
        if (f == 0 || f == n) {
            const float sgn = f == 0 ? 1.0f : -1.0f;
            const float vn = sgn * v[ix(i, f, P)];
            const float php = phi[ix(i, cell, P)];
            return (vn * php) / h;   // zero-gradient side: advective flux only
        }
        const int Bc = f, Tc = f + 1;
        const float vf = v[ix(i, f, P)];
        int C, Dn, Uj;
        if (vf >= 0.0f) { C = Bc; Dn = Tc; Uj = Bc - 1; }
        else { C = Tc; Dn = Bc; Uj = Tc + 1; }
        const bool has_u = 1 <= Uj && Uj <= n;
        const float phu = has_u ? phi[ix(i, Uj, P)] : 0.0f;
        const float phf = tvd_face(phu, phi[ix(i, C, P)], phi[ix(i, Dn, P)], has_u);
        const float df = D + D > 0.0f ? 2.0f * D * D / (D + D) : 0.0f;
        const float Fy = vf * phf - df * (phi[ix(i, Tc, P)] - phi[ix(i, Bc, P)]) / h;
        return cell == Bc ? -(Fy / h) : Fy / h;
    };
    float rhs = xterm(i - 1, i);
    rhs += xterm(i, i);
    rhs += yterm(j - 1, j);
    rhs += yterm(j, j);
    out[ix(i, j, P)] = phi[ix(i, j, P)] + (float)c->st[b].dt * rhs;
}

// temperature ghosts (models.py:139-183 for these codes), one work-group per instance
__global__ void __launch_bounds__(256) k_e_tghost(const EnsClass* __restrict__ c) {
    if (c->ctl->stall) return;
    const int par = c->ctl->par;
    const int b = blockIdx.x, n = c->n, P = c->P;
    if (c->kappa[b] <= 0.0) return;
    float* T = c->T[par ^ 1] + b * c->S;
    for (int k = 1 + threadIdx.x; k <= n; k += blockDim.x) {
        T[ix(0, k, P)] = 2.0f * 0.5f - T[ix(1, k, P)];
        T[ix(n + 1, k, P)] = 2.0f * -0.5f - T[ix(n, k, P)];
        T[ix(k, 0, P)] = T[ix(k, 1, P)];
        T[ix(k, n + 1, P)] = T[ix(k, n, P)];
    }
}

// monitors: partial max |div u| and sum of cell kinetic energy
__global__ void __launch_bounds__(256) k_e_record(const EnsClass* __restrict__ c) {
    if (c->ctl->stall) return;
    const int b = blockIdx.y, n = c->n, P = c->P;
    const long o = b * c->S;
    const float* u = c->u + o;
    const float* v = c->v + o;
    const double h = c->hd;
    int i0, i1;
    ens_band(1, n + 1, blockIdx.x, ENS_BANDS, i0, i1);
    double mx[1] = {0.0}, s[1] = {0.0};
    for (int i = i0; i < i1; ++i)
        for (int j = 1 + threadIdx.x; j <= n; j += blockDim.x) {
            const double div = ((double)u[ix(i, j, P)] - u[ix(i - 1, j, P)]) / h +
                               ((double)v[ix(i, j, P)] - v[ix(i, j - 1, P)]) / h;
            mx[0] = nanmax(mx[0], fabs(div));
            const double uc = 0.5 * ((double)u[ix(i - 1, j, P)] + u[ix(i, j, P)]);
            const double vc = 0.5 * ((double)v[ix(i, j - 1, P)] + v[ix(i, j, P)]);
            s[0] += uc * uc + vc * vc;
        }
    ens_partial_max<1>(mx, ens_part(c, ENS_P_C, b), blockIdx.x);
    ens_partial<1>(s, ens_part(c, ENS_P_D, b), blockIdx.x);
}

// per instance: monitors and the state advance (solver.py:378-401)
__global__ void __launch_bounds__(256) k_e_advance(const EnsClass* __restrict__ c) {
    if (c->ctl->stall) return;
    for (int b = threadIdx.x; b < c->B; b += blockDim.x) {
        EnsState& st = c->st[b];
        double mx = 0.0, ke = 0.0, res = 0.0;
        const double* pc = ens_part(c, ENS_P_C, b);
        const double* pd = ens_part(c, ENS_P_D, b);
        const double* pb = ens_part(c, ENS_P_B, b);
        for (int k = 0; k < ENS_BANDS; ++k) {
            mx = nanmax(mx, pc[k * ENS_NV]);
            ke += pd[k * ENS_NV];
            res = nanmax(res, pb[k * ENS_NV]);
        }
        const double dt = st.dt;
        st.divmax = mx;
        st.ke = 0.5 * ke * c->hd * c->hd;
        st.residual = res / dt;
        st.has_old = 1;
        st.dt_old = dt;
        st.t += dt;
        st.nstep += 1;
    }
}

// after a step's solver iterations (one work-group): whether every solve of
// every class has finished -- then the step's lockstep iterations (the most
// any instance took) are counted and the post-solve kernels run -- or the
// step stalls until the next launch continues its iterations
__global__ void __launch_bounds__(256) k_e_cap(const EnsClass* __restrict__ d, int ncls) {
    __shared__ int s_left, s_max;
    if (threadIdx.x == 0) {
        s_left = 0;
        s_max = 0;
    }
    __syncthreads();
    for (int k = 0; k < ncls; ++k) {
        const EnsClass* c = d + k;
        if (threadIdx.x == 0) atomicAdd(&s_left, *c->notdone);
        for (int b = threadIdx.x; b < c->B; b += blockDim.x) atomicMax(&s_max, c->st[b].k);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        EnsCtl* ctl = d->ctl;
        if (s_left > 0) {
            ctl->stall = 1;
            ctl->stalls += 1;
        } else {
            ctl->stall = 0;
            ctl->iters += s_max;
            ctl->last_max = s_max;
            ctl->max_call = max(ctl->max_call, s_max);
        }
    }
}

// end of a completed step: the next step's buffer parity and history
__global__ void k_e_flip(EnsCtl* ctl) {
    if (threadIdx.x == 0 && !ctl->stall) {
        ctl->par ^= 1;
        ctl->hist = ctl->x0ext;
    }
}
