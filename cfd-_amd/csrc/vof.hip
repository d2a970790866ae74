// Two-phase flow with the volume-of-fluid method (cfd2d/vof.py, interFoam's
// approach): the phase fraction alpha is transported by a bounded van Leer
// scheme plus the interface-compression flux u_r alpha (1 - alpha); density
// and dynamic viscosity are mixed from alpha; surface tension enters as the
// CSF body force sigma kappa grad(alpha); the pressure equation becomes
// div((1/rho) grad p) = div(u*) / dt (pressure.py:187-244, solver.py:267-273).
//   k_vof_kappa       curvature kappa = -div(n) from the smoothed alpha
//                     (vof.py:217-234), before the predictor, which adds the
//                     CSF force divided by the face density
//   k_project_var     cooperative Jacobi-PCG on the variable-coefficient
//                     system and the correction u = F - dt (1/rho_f) grad p
//   k_vof_nsub        sub-step count of the alpha transport (vof.py:243-248),
//                     a grid-wide maximum finished by the last block
//   k_vof_advect      one alpha step in gather form: every cell adds its
//                     faces' fluxes in the order the reference's face loops
//                     reach it, then clips to [0, 1] (vof.py:34-133)
//   k_vof_ghosts,     ghost cells, then solid cells and the properties
//   k_vof_solids      rho, mu (vof.py:182-210)
//   k_vof_monitors    phase volume, front position, column height
//                     (vof.py:264-281)
// Every expression keeps the reference's operand order.
#pragma clang fp contract(off)
#include "cfd_kernels.h"

struct VofArgs {
    const double* a;         // alpha (input of the kernel)
    double* out;             // alpha written
    const double* u;
    const double* v;
    const uint8_t* fl;
    double* rho;
    double* mu;
    double* kappa;
    DevState* st;
    double rho1, rho2, mu1, mu2;
    double c_alpha, cfl, a_in;
    int sub;                 // k_vof_advect: this launch's sub-step index
    int nsub_max;            // sub-step launches per step
    double* part;            // per-block partials of the monitors (3 per block)
    unsigned* ticket;        // [0] k_vof_nsub, [1] k_vof_monitors
    unsigned long long* acc; // max |u|, max |v| (bits of non-negative doubles)
};

// ---- surface tension: curvature of the smoothed alpha ----------------------
// smoothed alpha of an inner cell ((4 a + E + W + N + S) / 8), alpha itself
// on the ghost ring
__device__ __forceinline__ double vof_smooth(const double* a, const Geo& g, int i, int j) {
    const int P = g.pc;
    if (i < 1 || i > g.nx || j < 1 || j > g.ny) return a[ix(i, j, P)];
    const long c = ix(i, j, P);
    return (4 * a[c] + a[c + P] + a[c - P] + a[c + 1] + a[c - 1]) / 8.0;
}

// unit normal at node (i, j) (between cells i-1..i, j-1..j of the padded grid)
__device__ __forceinline__ void vof_normal(const double* a, const Geo& g, int i, int j,
                                           double& nxn, double& nyn) {
    const double s11 = vof_smooth(a, g, i, j), s01 = vof_smooth(a, g, i - 1, j);
    const double s10 = vof_smooth(a, g, i, j - 1), s00 = vof_smooth(a, g, i - 1, j - 1);
    const double gx = 0.5 * ((s11 - s01) + (s10 - s00)) / g.dx;
    const double gy = 0.5 * ((s11 - s10) + (s01 - s00)) / g.dy;
    const double gm = sqrt(gx * gx + gy * gy) + 1e-8 / g.dx;
    nxn = gx / gm;
    nyn = gy / gm;
}

__global__ void __launch_bounds__(256) k_vof_kappa(VofArgs a, Geo g) {
    const int2 rc_ = xcd_rowgrid();
    const int j = rc_.x * blockDim.x + threadIdx.x + 1;
    const int i = rc_.y + 1;
    if (j > g.ny) return;
    if (a.st->done) return;
    double nx11, ny11, nx01, ny01, nx10, ny10, nx00, ny00;
    vof_normal(a.a, g, i + 1, j + 1, nx11, ny11);
    vof_normal(a.a, g, i, j + 1, nx01, ny01);
    vof_normal(a.a, g, i + 1, j, nx10, ny10);
    vof_normal(a.a, g, i, j, nx00, ny00);
    a.kappa[ix(i, j, g.pc)] = -(0.5 * ((nx11 - nx01) + (nx10 - nx00)) / g.dx +
                                0.5 * ((ny11 - ny10) + (ny01 - ny00)) / g.dy);
}

// ---- alpha transport ---------------------------------------------------------
__device__ __forceinline__ double vof_vl(double pu, double pc, double pd, bool has_u) {
    if (!has_u) return pc;
    const double den = pd - pc;
    if (fabs(den) < 1e-300) return pc;
    const double r = (pc - pu) / den;
    const double psi = r <= 0.0 ? 0.0 : 2.0 * r / (1.0 + r);
    return pc + 0.5 * psi * den;
}

// flux through the inner x face between cells L and R (R wraps to 1 for the
// periodic face nx), row j
__device__ double vof_xflux(const VofArgs& a, const Geo& g, int L, int R, int j) {
    const int nx = g.nx, ny = g.ny, P = g.pc;
    const double* al = a.a;
    const double uf = a.u[ix(L, j, g.pu)];
    int C, D, Ui;
    if (uf >= 0.0) {
        C = L; D = R; Ui = L - 1;
        if (Ui < 1) Ui = g.periodic ? nx : -1;
    } else {
        C = R; D = L; Ui = R + 1;
        if (Ui > nx) Ui = g.periodic ? 1 : -1;
    }
    const bool has_u = Ui >= 1 && (a.fl[ix(Ui, j, P)] & 1);
    const double af = vof_vl(has_u ? al[ix(Ui, j, P)] : 0.0, al[ix(C, j, P)], al[ix(D, j, P)], has_u);
    double flux = uf * af;
    const double gx = (al[ix(R, j, P)] - al[ix(L, j, P)]) / g.dx;
    const int jn = j < ny ? j + 1 : j, js = j > 1 ? j - 1 : j;
    const double gy = 0.25 * ((al[ix(L, jn, P)] - al[ix(L, js, P)]) +
                              (al[ix(R, jn, P)] - al[ix(R, js, P)])) / g.dy;
    const double gm = sqrt(gx * gx + gy * gy);
    if (gm > 1e-12 && a.c_alpha > 0.0) {
        const double ur = a.c_alpha * fabs(uf) * gx / gm;
        const double ac = ur >= 0.0 ? al[ix(L, j, P)] : al[ix(R, j, P)];
        const double ad = ur >= 0.0 ? al[ix(R, j, P)] : al[ix(L, j, P)];
        flux += ur * ac * (1.0 - ad);
    }
    return flux;
}

// flux through the inner y face between cells B = j and T = j + 1, column i
__device__ double vof_yflux(const VofArgs& a, const Geo& g, int i, int B) {
    const int nx = g.nx, ny = g.ny, P = g.pc;
    const double* al = a.a;
    const int T = B + 1;
    const double vf = a.v[ix(i, B, g.pv)];
    int C, D, Uj;
    if (vf >= 0.0) { C = B; D = T; Uj = B - 1; }
    else { C = T; D = B; Uj = T + 1; }
    const bool has_u = Uj >= 1 && Uj <= ny && (a.fl[ix(i, Uj, P)] & 1);
    const double af = vof_vl(has_u ? al[ix(i, Uj, P)] : 0.0, al[ix(i, C, P)], al[ix(i, D, P)], has_u);
    double flux = vf * af;
    const double gy = (al[ix(i, T, P)] - al[ix(i, B, P)]) / g.dy;
    int ie = (i < nx || g.periodic) ? i + 1 : i;
    int iw = (i > 1 || g.periodic) ? i - 1 : i;
    if (ie > nx) ie = 1;
    if (iw < 1) iw = nx;
    const double gx = 0.25 * ((al[ix(ie, B, P)] - al[ix(iw, B, P)]) +
                              (al[ix(ie, T, P)] - al[ix(iw, T, P)])) / g.dx;
    const double gm = sqrt(gx * gx + gy * gy);
    if (gm > 1e-12 && a.c_alpha > 0.0) {
        const double ur = a.c_alpha * fabs(vf) * gy / gm;
        const double ac = ur >= 0.0 ? al[ix(i, B, P)] : al[ix(i, T, P)];
        const double ad = ur >= 0.0 ? al[ix(i, T, P)] : al[ix(i, B, P)];
        flux += ur * ac * (1.0 - ad);
    }
    return flux;
}

// sub-step count from the corrected velocity: max |u| / dx, |v| / dy over all
// faces (one work-group; tiny next to the step)
__global__ void __launch_bounds__(256) k_vof_nsub(VofArgs a, Geo g) {
    __shared__ double lds[2 * 16];
    if (a.st->done) return;
    const long nu = (long)(g.nx + 1) * g.pu, nv = (long)(g.nx + 2) * g.pv;
    const long k0 = (long)blockIdx.x * blockDim.x + threadIdx.x, ks = (long)gridDim.x * blockDim.x;
    double m[2] = {0.0, 0.0};
    for (long k = k0; k < nu; k += ks)
        if (k % g.pu <= g.ny + 1) m[0] = nanmax(m[0], fabs(a.u[k]));
    for (long k = k0; k < nv; k += ks)
        if (k % g.pv <= g.ny) m[1] = nanmax(m[1], fabs(a.v[k]));
    block_max<2>(m, lds);
    __shared__ int s_last;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < 2; ++q)
            __hip_atomic_fetch_max(&a.acc[q], (unsigned long long)__double_as_longlong(m[q]),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned tk =
            __hip_atomic_fetch_add(&a.ticket[0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        s_last = tk == gridDim.x - 1;
    }
    __syncthreads();
    if (!s_last || threadIdx.x != 0) return;
    double mx[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        mx[q] = __longlong_as_double(
            (long long)__hip_atomic_load(&a.acc[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        a.acc[q] = 0ull;
    }
    a.ticket[0] = 0u;
    const double rate = fmax(fmax(mx[0] / g.dx, mx[1] / g.dy), 1e-12);
    const double dt = a.st->dt;
    const double nsub = fmax(1.0, ceil(dt * rate / a.cfl));
    a.st->vof_nsub = nsub;
    if (nsub > a.nsub_max) {   // more sub-steps than launched: stop as unstable
        a.st->error = 1;
        a.st->done = 1;
    }
}

// one sub-step; skipped beyond the sub-step count (the host issues a fixed
// number of launches)
__global__ void __launch_bounds__(256) k_vof_advect(VofArgs a, Geo g) {
    const int2 rc_ = xcd_rowgrid();
    const int j = rc_.x * blockDim.x + threadIdx.x;
    const int i = rc_.y;
    const int nx = g.nx, ny = g.ny, P = g.pc;
    if (j > ny + 1) return;
    if (a.st->done) return;
    const long c = ix(i, j, P);
    double val = a.a[c];
    const bool inner = i >= 1 && i <= nx && j >= 1 && j <= ny;
    // beyond the sub-step count the field is passed on unchanged, so that it
    // ends in the same buffer whatever the count
    if (!inner || !(a.fl[c] & 1) || (double)a.sub >= a.st->vof_nsub) {
        a.out[c] = val;
        return;
    }
    const double h = a.st->dt / a.st->vof_nsub;
    auto fluid = [&](int ci, int cj) -> bool { return a.fl[ix(ci, cj, P)] & 1; };
    // x faces in the reference's loop order: the west face, then the east face
    // (the periodic face nx reaches cell 1 last)
    if (i == 1) {
        if (!g.periodic) {
            const double into = a.u[ix(0, j, g.pu)];
            const double ab = into > 0 ? a.a_in : val;
            val += h * into * ab / g.dx;
        }
    } else if (fluid(i - 1, j)) {
        val += h * vof_xflux(a, g, i - 1, i, j) / g.dx;
    }
    if (i == nx) {
        if (!g.periodic) {
            const double into = -a.u[ix(nx, j, g.pu)];
            const double ab = into > 0 ? a.a_in : a.a[c];
            val += h * into * ab / g.dx;
        } else if (fluid(1, j)) {
            val -= h * vof_xflux(a, g, nx, 1, j) / g.dx;
        }
    } else if (fluid(i + 1, j)) {
        val -= h * vof_xflux(a, g, i, i + 1, j) / g.dx;
    }
    if (i == 1 && g.periodic && fluid(nx, j)) val += h * vof_xflux(a, g, nx, 1, j) / g.dx;
    // y faces: south, then north
    if (j == 1) {
        const double into = a.v[ix(i, 0, g.pv)];
        const double ab = into > 0 ? a.a_in : a.a[c];
        val += h * into * ab / g.dy;
    } else if (fluid(i, j - 1)) {
        val += h * vof_yflux(a, g, i, j - 1) / g.dy;
    }
    if (j == ny) {
        const double into = -a.v[ix(i, ny, g.pv)];
        const double ab = into > 0 ? a.a_in : a.a[c];
        val += h * into * ab / g.dy;
    } else if (fluid(i, j + 1)) {
        val -= h * vof_yflux(a, g, i, j) / g.dy;
    }
    if (val < 0.0) val = 0.0;
    else if (val > 1.0) val = 1.0;
    a.out[c] = val;
}

// ghost ring of alpha: the columns over the full height (periodic copies or
// the neighbour), then the rows over the full width (vof.py:182-192)
__global__ void __launch_bounds__(256) k_vof_ghosts(VofArgs a, Geo g) {
    const int nx = g.nx, ny = g.ny, P = g.pc;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (a.st->done || (a.sub >= 0 && (double)a.sub >= a.st->vof_nsub)) return;
    double* al = a.out;
    auto col = [&](int ci, int cj) -> double {
        if (g.periodic) return al[ix(ci == 0 ? nx : 1, cj, P)];
        return al[ix(ci == 0 ? 1 : nx, cj, P)];
    };
    int i, j;
    if (t < 2 * (nx + 2)) {
        i = t % (nx + 2);
        j = t < nx + 2 ? 0 : ny + 1;
        const int jn = j == 0 ? 1 : ny;
        al[ix(i, j, P)] = (i >= 1 && i <= nx) ? al[ix(i, jn, P)] : col(i, jn);
    } else if (t < 2 * (nx + 2) + 2 * ny) {
        const int r = t - 2 * (nx + 2);
        j = 1 + r % ny;
        i = r < ny ? 0 : nx + 1;
        al[ix(i, j, P)] = col(i, j);
    }
}

// solid cells: mean of the fluid neighbours (0 without any); then rho and mu
// of every cell from the final alpha (vof.py:194-210)
__global__ void __launch_bounds__(256) k_vof_solids(VofArgs a, Geo g) {
    const int2 rc_ = xcd_rowgrid();
    const int j = rc_.x * blockDim.x + threadIdx.x;
    const int i = rc_.y;
    if (j > g.ny + 1) return;
    if (a.st->done || (a.sub >= 0 && (double)a.sub >= a.st->vof_nsub)) return;
    const int P = g.pc;
    const long c = ix(i, j, P);
    double x = a.out[c];
    const bool inner = i >= 1 && i <= g.nx && j >= 1 && j <= g.ny;
    // only fluid neighbours are read and only solid cells written, so the
    // update can be in place
    if (inner && !(a.fl[c] & 1)) {
        auto fv = [&](long cc) -> double { return (a.fl[cc] & 1) ? a.out[cc] : 0.0; };
        auto fc = [&](long cc) -> double { return (a.fl[cc] & 1) ? 1.0 : 0.0; };
        const double s = fv(c + P) + fv(c - P) + fv(c + 1) + fv(c - 1);
        const double n = fc(c + P) + fc(c - P) + fc(c + 1) + fc(c - 1);
        x = n > 0 ? s / fmax(n, 1.0) : 0.0;
        a.out[c] = x;
    }
    if (a.rho) {
        a.rho[c] = x * a.rho1 + (1 - x) * a.rho2;
        a.mu[c] = x * a.mu1 + (1 - x) * a.mu2;
    }
}

// phase volume sum over fluid cells, front (last cell of the bottom row with
// alpha > 0.5) and height (last cell of the left column): one work-group
__global__ void __launch_bounds__(256) k_vof_monitors(VofArgs a, Geo g) {
    __shared__ double lds[3 * 16];
    if (a.st->done) return;
    const int nx = g.nx, ny = g.ny, P = g.pc;
    const int r0 = 1 + (int)(((long)nx * blockIdx.x) / gridDim.x);
    const int r1 = 1 + (int)(((long)nx * (blockIdx.x + 1)) / gridDim.x);
    double s[1] = {0.0};
    double m[2] = {-1.0, -1.0};
    for (int i = r0; i < r1; ++i)
        for (int j = 1 + threadIdx.x; j <= ny; j += blockDim.x) {
            const long c = ix(i, j, P);
            if (a.fl[c] & 1) s[0] += a.out[c];
            if (j == 1 && a.out[c] > 0.5) m[0] = fmax(m[0], (double)i);
            if (i == 1 && a.out[c] > 0.5) m[1] = fmax(m[1], (double)j);
        }
    block_sum<1>(s, lds);
    block_max<2>(m, lds);
    __shared__ int s_last;
    if (threadIdx.x == 0) {
        a.part[3 * blockIdx.x] = s[0];
        a.part[3 * blockIdx.x + 1] = m[0];
        a.part[3 * blockIdx.x + 2] = m[1];
        const unsigned tk =
            __hip_atomic_fetch_add(&a.ticket[1], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        s_last = tk == gridDim.x - 1;
    }
    __syncthreads();
    if (!s_last) return;
    // the last block adds the partials in a fixed order
    double t[1] = {0.0};
    double n[2] = {-1.0, -1.0};
    for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) {
        t[0] += __hip_atomic_load(&a.part[3 * b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        n[0] = fmax(n[0], __hip_atomic_load(&a.part[3 * b + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        n[1] = fmax(n[1], __hip_atomic_load(&a.part[3 * b + 2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    }
    block_sum<1>(t, lds);
    block_max<2>(n, lds);
    if (threadIdx.x != 0) return;
    a.ticket[1] = 0u;
    a.st->vof_sum = t[0];
    a.st->vof_front = n[0];
    a.st->vof_height = n[1];
}

// ---- variable-density pressure solve ----------------------------------------
// Aggregation multigrid for the variable-coefficient system, rebuilt every
// solve: level l+1 is 0.5 P^T A_l P for 2x2-cell aggregates (coupling of two
// aggregates = half the fine couplings between them; diagonal = half the
// fine diagonals minus twice the couplings inside), damped Jacobi (0.8) two
// sweeps before and after, an exact solve on the coarsest level of at most
// VMG_COARSE cells.  The V-cycle is symmetric, so it preconditions the CG
// directly.
constexpr int VMG_MAX = 12;
constexpr int VMG_COARSE = 64;
struct VLev {
    int nx, ny, P;           // cells, pitch of the padded (nx+2) x (ny+2) arrays
    double *cE, *cN, *diag;  // operator
    double *x, *t, *b;       // two iterates and the right-hand side
};
struct VMG {
    int nlev;                // 0: plain Jacobi-PCG
    VLev lev[VMG_MAX];       // lev[0] is the fine grid (x, t work vectors, b = r)
};

struct VarProjArgs {
    ProjArgs p;              // F, G, u, v, p, code, comp, comp_count, ncomp, x, b, r, q, d0,
                             // sync, st, rtol, maxiter, dt_override
    const double* rho;
    double* cE;              // coupling to the east / north neighbour (0 where none)
    double* cN;
    double* diag;
    const int* pinned;       // cell offsets pinned in pure-Neumann components
    int npinned;
    VMG mg;
};

__device__ __forceinline__ double bx_face(const double* rho, const Geo& g, int f, int j) {
    // x face f between padded cells f and f+1; the periodic face nx couples
    // cells nx and 1 (solver.py:269-272)
    const int P = g.pc;
    if (g.periodic && f == g.nx) return 2.0 / (rho[ix(g.nx, j, P)] + rho[ix(1, j, P)]);
    return 2.0 / (rho[ix(f, j, P)] + rho[ix(f + 1, j, P)]);
}

__device__ __forceinline__ double by_face(const double* rho, const Geo& g, int i, int f) {
    return 2.0 / (rho[ix(i, f, g.pc)] + rho[ix(i, f + 1, g.pc)]);
}

__global__ void __launch_bounds__(PROJ_BS) k_project_var(VarProjArgs va, Geo g) {
    __shared__ double lds[CFD_MAX_COMP * (PROJ_BS / 64) + 16];
    __shared__ double cinv[VMG_COARSE * VMG_COARSE];   // inverse of the coarsest operator
    __shared__ double cb[VMG_COARSE];
    const ProjArgs& a = va.p;
    DevState* st = a.st;
    if (st->done) return;
    unsigned epoch = 0;
    const int nx = g.nx, ny = g.ny, pc = g.pc;
    const int periodic = g.periodic;
    const double dx2 = 1.0 / (g.dx * g.dx), dy2 = 1.0 / (g.dy * g.dy);
    const double dt = a.dt_override > 0.0 ? a.dt_override : st->dt;
    const int i0 = 1 + (int)(((long)nx * blockIdx.x) / gridDim.x);
    const int i1 = 1 + (int)(((long)nx * (blockIdx.x + 1)) / gridDim.x);
    const int tid = threadIdx.x, BS = blockDim.x;
    const bool lead = blockIdx.x == 0 && tid == 0;
    const GridSync gs{a.sync, a.sync + (long)gridDim.x * CFD_SLOT, a.spin};
    auto fluid = [&](int i, int j) -> bool { return a.code[ix(i, j, pc)] & 63; };
#define FOR_CELLS                          \
    for (int i = i0; i < i1; ++i)          \
        for (int j = 1 + tid; j <= ny; j += BS)

    // ---- couplings, right-hand side, the largest x-face coefficient ----
    double cs[CFD_MAX_COMP];
#pragma unroll
    for (int k = 0; k < CFD_MAX_COMP; ++k) cs[k] = 0.0;
    double bxm[1] = {0.0};
    FOR_CELLS {
        const long c = ix(i, j, pc);
        const bool f = fluid(i, j);
        double ce = 0.0, cn = 0.0;
        if (f) {
            if (i < nx ? fluid(i + 1, j) : (periodic && fluid(1, j))) ce = bx_face(va.rho, g, i, j) * dx2;
            if (j < ny && fluid(i, j + 1)) cn = by_face(va.rho, g, i, j) * dy2;
        }
        va.cE[c] = ce;
        va.cN[c] = cn;
        bxm[0] = fmax(bxm[0], bx_face(va.rho, g, i, j));
        if (i == 1) bxm[0] = fmax(bxm[0], bx_face(va.rho, g, 0, j));
        if (!f) continue;
        const double div = (a.F[ix(i, j, g.pu)] - a.F[ix(i - 1, j, g.pu)]) / g.dx +
                           (a.G[ix(i, j, g.pv)] - a.G[ix(i, j - 1, g.pv)]) / g.dy;
        const double rhs = div / dt;
        a.b[c] = rhs;
        const int cm = a.comp[c];
#pragma unroll
        for (int k = 0; k < CFD_MAX_COMP; ++k) cs[k] += (cm == k + 1) ? rhs : 0.0;
    }
    grid_max<1>(bxm, gs, lds, epoch);   // also orders the couplings before their use
    double mean[CFD_MAX_COMP];
    if (a.ncomp > 0) {
        grid_sum<CFD_MAX_COMP>(cs, gs, lds, epoch);
#pragma unroll
        for (int k = 0; k < CFD_MAX_COMP; ++k)
            mean[k] = k < a.ncomp ? cs[k] / a.comp_count[k] : 0.0;
    } else {
#pragma unroll
        for (int k = 0; k < CFD_MAX_COMP; ++k) mean[k] = 0.0;
    }
    const double pin_add = dx2 * fmax(bxm[0], 1e-30);

    // ---- diagonal in the reference's order of additions, b = -(rhs - mean),
    //      r = b - A x, z = r / diag ----
    auto cW = [&](int i, int j) -> double {
        return i > 1 ? va.cE[ix(i - 1, j, pc)] : 0.0;
    };
    auto apply = [&](const double* d, int i, int j, double dg) -> double {
        const long c = ix(i, j, pc);
        const int iE = (periodic && i == nx) ? 1 : i + 1;
        const int iW = (periodic && i == 1) ? nx : i - 1;
        const double ce = va.cE[c];
        const double cw = (periodic && i == 1) ? va.cE[ix(nx, j, pc)] : cW(i, j);
        const double cn = va.cN[c], csv = j > 1 ? va.cN[c - 1] : 0.0;
        return dg * d[c] - ce * d[ix(iE, j, pc)] - cw * d[ix(iW, j, pc)] - cn * d[c + 1] -
               csv * d[c - 1];
    };
    double s3[3] = {0.0, 0.0, 0.0};
    FOR_CELLS {
        const long c = ix(i, j, pc);
        if (!fluid(i, j)) continue;
        const int code = a.code[c];
        double dg = 0.0;
        if (i < nx) dg += va.cE[c];
        dg += cW(i, j);
        dg += va.cN[c];
        if (j > 1) dg += va.cN[c - 1];
        if (periodic && i == nx) dg += va.cE[c];
        if (periodic && i == 1) dg += va.cE[ix(nx, j, pc)];
        if (code & 64) {
            if (i == 1) dg += 2 * dx2 * bx_face(va.rho, g, 0, j);
            if (i == nx) dg += 2 * dx2 * bx_face(va.rho, g, nx, j);
        }
        if (code & 128) {
            if (j == 1) dg += 2 * dy2 * by_face(va.rho, g, i, 0);
            if (j == ny) dg += 2 * dy2 * by_face(va.rho, g, i, ny);
        }
        for (int k = 0; k < va.npinned; ++k)
            if (va.pinned[k] == c) dg += pin_add;
        va.diag[c] = dg;
        double bv = a.b[c];
        const int cm = a.comp[c];
        if (cm) bv = bv - sel_comp(mean, cm);
        bv = -bv;
        a.b[c] = bv;
        const double rv = bv - apply(a.x, i, j, dg);
        a.r[c] = rv;
        const double z = rv / dg;
        s3[0] += bv * bv;
        s3[1] += rv * rv;
        s3[2] += rv * z;
    }
    grid_sum<3>(s3, gs, lds, epoch);
    const double atol = a.rtol * sqrt(s3[0]);
    double rr = s3[1], rz = s3[2];
    int iters = 0, info = 0;
    if (s3[0] == 0.0) {
        FOR_CELLS {
            if (fluid(i, j)) a.x[ix(i, j, pc)] = 0.0;
        }
        rr = 0.0;
    } else if (va.mg.nlev >= 2) {
        // ---- multigrid-preconditioned CG ----
        const VMG& mg = va.mg;
        const int L = mg.nlev;
        constexpr double om = 0.8;
        // a level of at most four cells per thread of one block is handled by
        // block 0 alone, behind block barriers instead of grid barriers
        auto is_small = [&](int l) -> bool { return (long)mg.lev[l].nx * mg.lev[l].ny <= 4L * BS; };
        auto lsync = [&](bool small) {
            if (small) __syncthreads();
            else grid_sync(gs, lds, epoch);
        };
        // body(i, j, c) for the cells of level l; consecutive threads take
        // consecutive j, the contiguous index
        auto each = [&](int l, auto&& body) {
            const VLev& V = mg.lev[l];
            const int n = V.nx * V.ny, vny = V.ny;
            int q = tid, qs = BS;
            if (is_small(l)) {
                if (blockIdx.x != 0) return;
            } else {
                q += blockIdx.x * BS;
                qs *= gridDim.x;
            }
            for (; q < n; q += qs) {
                const int i = q / vny, j = q - i * vny;
                body(i + 1, j + 1, ix(i + 1, j + 1, V.P));
            }
        };
        // (A xv)(i, j) on level V, xv(i, j, c) giving the iterate; couplings
        // that are zero do not read their neighbour
        auto Aop = [&](const VLev& V, int i, int j, long c, auto&& xv) -> double {
            const int iE = (periodic && i == V.nx) ? 1 : i + 1;
            const int iW = (periodic && i == 1) ? V.nx : i - 1;
            const double ce = V.cE[c];
            const double cw = (periodic && i == 1) ? V.cE[ix(V.nx, j, V.P)]
                                                   : (i > 1 ? V.cE[c - V.P] : 0.0);
            const double cn = V.cN[c], cs = j > 1 ? V.cN[c - 1] : 0.0;
            double r = V.diag[c] * xv(i, j, c);
            if (ce != 0.0) r -= ce * xv(iE, j, ix(iE, j, V.P));
            if (cw != 0.0) r -= cw * xv(iW, j, ix(iW, j, V.P));
            if (cn != 0.0) r -= cn * xv(i, j + 1, c + 1);
            if (cs != 0.0) r -= cs * xv(i, j - 1, c - 1);
            return r;
        };
        // coarse operators
        for (int l = 0; l + 1 < L; ++l) {
            const VLev& F = mg.lev[l];
            const VLev& C = mg.lev[l + 1];
            each(l + 1, [&](int i, int j, long c) {
                const int i0 = 2 * i - 1, j0 = 2 * j - 1;
                const bool hi = 2 * i <= F.nx, hj = 2 * j <= F.ny;
                const int i1 = hi ? i0 + 1 : i0, j1 = hj ? j0 + 1 : j0;
                auto fe = [&](int fi, int fj) { return F.cE[ix(fi, fj, F.P)]; };
                auto fn = [&](int fi, int fj) { return F.cN[ix(fi, fj, F.P)]; };
                auto fd = [&](int fi, int fj) { return F.diag[ix(fi, fj, F.P)]; };
                double ds = fd(i0, j0), inner = 0.0;
                if (hi) { ds += fd(i1, j0); inner += fe(i0, j0); }
                if (hj) { ds += fd(i0, j1); inner += fn(i0, j0); }
                if (hi && hj) { ds += fd(i1, j1); inner += fe(i0, j1) + fn(i1, j0); }
                C.diag[c] = 0.5 * (ds - 2.0 * inner);
                C.cE[c] = 0.5 * (fe(i1, j0) + (hj ? fe(i1, j1) : 0.0));
                C.cN[c] = 0.5 * (fn(i0, j1) + (hi ? fn(i1, j1) : 0.0));
            });
            lsync(is_small(l + 1) && (l + 2 >= L || is_small(l + 2)));
        }
        grid_sync(gs, lds, epoch);
        // the coarsest level (at most VMG_COARSE cells) is solved exactly: its
        // dense matrix, in the cell order q = (i-1) ny + j-1, is inverted in
        // LDS by Gauss-Jordan elimination without pivoting (the matrix is
        // symmetric positive definite; all-solid aggregates get a unit row)
        const VLev& Vc = mg.lev[L - 1];
        const int nco = Vc.nx * Vc.ny;
        if (blockIdx.x == 0) {
            for (int e = tid; e < VMG_COARSE * VMG_COARSE; e += BS) cinv[e] = 0.0;
            __syncthreads();
            for (int q = tid; q < nco; q += BS) {
                const int i = 1 + q / Vc.ny, j = 1 + q % Vc.ny;
                const long c = ix(i, j, Vc.P);
                double* row = cinv + q * VMG_COARSE;
                if (!(Vc.diag[c] > 0.0)) {
                    row[q] = 1.0;
                    continue;
                }
                const int iE = (periodic && i == Vc.nx) ? 1 : i + 1;
                const int iW = (periodic && i == 1) ? Vc.nx : i - 1;
                const double cw = (periodic && i == 1) ? Vc.cE[ix(Vc.nx, j, Vc.P)]
                                                       : (i > 1 ? Vc.cE[c - Vc.P] : 0.0);
                row[q] += Vc.diag[c];
                if (Vc.cE[c] != 0.0) row[(iE - 1) * Vc.ny + j - 1] -= Vc.cE[c];
                if (cw != 0.0) row[(iW - 1) * Vc.ny + j - 1] -= cw;
                if (Vc.cN[c] != 0.0) row[q + 1] -= Vc.cN[c];
                if (j > 1 && Vc.cN[c - 1] != 0.0) row[q - 1] -= Vc.cN[c - 1];
            }
            __syncthreads();
            for (int k = 0; k < nco; ++k) {
                const double piv = cinv[k * VMG_COARSE + k];
                for (int e = tid; e < nco * VMG_COARSE; e += BS) {
                    const int r = e / VMG_COARSE, s = e % VMG_COARSE;
                    if (r != k && s < nco && s != k)
                        cinv[e] -= cinv[r * VMG_COARSE + k] * cinv[k * VMG_COARSE + s] / piv;
                }
                __syncthreads();
                for (int e = tid; e < nco; e += BS) {
                    if (e == k) continue;
                    cinv[e * VMG_COARSE + k] = -cinv[e * VMG_COARSE + k] / piv;
                    cinv[k * VMG_COARSE + e] = cinv[k * VMG_COARSE + e] / piv;
                }
                if (tid == 0) cinv[k * VMG_COARSE + k] = 1.0 / piv;
                __syncthreads();
            }
        }
        // z = M r by one V-cycle from zero; z is lev[0].t and the return value
        // is this thread's share of r . z
        auto vcycle = [&]() -> double {
            for (int l = 0; l + 1 < L; ++l) {
                const VLev& V = mg.lev[l];
                const VLev& C = mg.lev[l + 1];
                if (l > 0) lsync(is_small(l));   // b of level l complete
                // two damped Jacobi sweeps from zero, the first formed for
                // the neighbours on the fly
                auto x1 = [&](int, int, long cc) -> double {
                    const double d = V.diag[cc];
                    return d > 0.0 ? om * V.b[cc] / d : 0.0;
                };
                each(l, [&](int i, int j, long c) {
                    const double d = V.diag[c];
                    V.t[c] = d > 0.0 ? x1(i, j, c) + om * (V.b[c] - Aop(V, i, j, c, x1)) / d : 0.0;
                });
                lsync(is_small(l) && is_small(l + 1));   // t of level l complete
                auto tv = [&](int, int, long cc) -> double { return V.t[cc]; };
                each(l + 1, [&](int i, int j, long c) {   // restriction of the residual
                    const int i0 = 2 * i - 1, j0 = 2 * j - 1;
                    double s = 0.0;
                    for (int di = 0; di < 2; ++di)
                        for (int dj = 0; dj < 2; ++dj) {
                            const int fi = i0 + di, fj = j0 + dj;
                            if (fi > V.nx || fj > V.ny) continue;
                            const long fc = ix(fi, fj, V.P);
                            if (V.diag[fc] > 0.0) s += V.b[fc] - Aop(V, fi, fj, fc, tv);
                        }
                    C.b[c] = s;
                });
            }
            lsync(true);   // b of the coarsest level complete (block 0 wrote it)
            if (blockIdx.x == 0) {
                for (int q = tid; q < nco; q += BS) cb[q] = Vc.b[ix(1 + q / Vc.ny, 1 + q % Vc.ny, Vc.P)];
                __syncthreads();
                for (int q = tid; q < nco; q += BS) {
                    double s = 0.0;
                    for (int k = 0; k < nco; ++k) s += cinv[k * VMG_COARSE + q] * cb[k];
                    const long c = ix(1 + q / Vc.ny, 1 + q % Vc.ny, Vc.P);
                    Vc.t[c] = Vc.diag[c] > 0.0 ? s : 0.0;
                }
            }
            double s1 = 0.0;
            for (int l = L - 2; l >= 0; --l) {
                const VLev& V = mg.lev[l];
                const VLev& C = mg.lev[l + 1];
                lsync(is_small(l) && is_small(l + 1));   // t of level l+1 complete
                // prolongation (add the aggregate's correction) fused with the
                // first of two damped Jacobi sweeps
                auto tp = [&](int ii, int jj, long cc) -> double {
                    return V.diag[cc] > 0.0 ? V.t[cc] + C.t[ix((ii + 1) / 2, (jj + 1) / 2, C.P)] : 0.0;
                };
                each(l, [&](int i, int j, long c) {
                    const double d = V.diag[c];
                    V.x[c] = d > 0.0 ? tp(i, j, c) + om * (V.b[c] - Aop(V, i, j, c, tp)) / d : 0.0;
                });
                lsync(is_small(l));
                auto xv = [&](int, int, long cc) -> double { return V.x[cc]; };
                each(l, [&](int i, int j, long c) {
                    const double d = V.diag[c];
                    const double z = d > 0.0 ? V.x[c] + om * (V.b[c] - Aop(V, i, j, c, xv)) / d : 0.0;
                    V.t[c] = z;
                    if (l == 0) s1 += V.b[c] * z;
                });
            }
            return s1;
        };
        const double* z = mg.lev[0].t;
        bool conv = false;
        double rz_prev = 1.0;
        rz = 0.0;
        for (int k = 0; k < a.maxiter; ++k) {
            if (sqrt(rr) < atol || !(rr == rr)) {
                conv = sqrt(rr) < atol;
                break;
            }
            double s1[1] = {vcycle()};
            grid_sum<1>(s1, gs, lds, epoch);   // also completes z
            rz_prev = rz;
            rz = s1[0];
            const double beta = k > 0 ? rz / rz_prev : 0.0;
            double* dnew = (k & 1) ? a.d1 : a.d0;
            const double* dold = (k & 1) ? a.d0 : a.d1;
            auto dirv = [&](long cc) -> double { return k > 0 ? z[cc] + beta * dold[cc] : z[cc]; };
            double dq[1] = {0.0};
            FOR_CELLS {
                const long c = ix(i, j, pc);
                if (!fluid(i, j)) continue;
                const int iE = (periodic && i == nx) ? 1 : i + 1;
                const int iW = (periodic && i == 1) ? nx : i - 1;
                const double ce = va.cE[c];
                const double cw = (periodic && i == 1) ? va.cE[ix(nx, j, pc)] : cW(i, j);
                const double cn = va.cN[c], csv = j > 1 ? va.cN[c - 1] : 0.0;
                const double dc = dirv(c);
                double q = va.diag[c] * dc;
                if (ce != 0.0) q -= ce * dirv(ix(iE, j, pc));
                if (cw != 0.0) q -= cw * dirv(ix(iW, j, pc));
                if (cn != 0.0) q -= cn * dirv(c + 1);
                if (csv != 0.0) q -= csv * dirv(c - 1);
                dnew[c] = dc;
                a.q[c] = q;
                dq[0] += dc * q;
            }
            grid_sum<1>(dq, gs, lds, epoch);
            const double alpha = rz / dq[0];
            double s2[1] = {0.0};
            FOR_CELLS {
                const long c = ix(i, j, pc);
                if (!fluid(i, j)) continue;
                a.x[c] += alpha * dnew[c];
                const double rv = a.r[c] - alpha * a.q[c];
                a.r[c] = rv;
                s2[0] += rv * rv;
            }
            grid_sum<1>(s2, gs, lds, epoch);
            rr = s2[0];
            ++iters;
        }
        if (!conv) info = iters > 0 ? iters : a.maxiter;
    } else {
        bool conv = false;
        double rz_prev = 1.0;
        for (int k = 0; k < a.maxiter; ++k) {
            if (sqrt(rr) < atol || !(rr == rr)) {
                conv = sqrt(rr) < atol;
                break;
            }
            // new direction d = z + beta d_old, formed for the neighbours too
            // (from their residual and old direction), so that the update and
            // the product q = A d need no grid barrier between them
            const double beta = k > 0 ? rz / rz_prev : 0.0;
            double* dnew = (k & 1) ? a.d1 : a.d0;
            const double* dold = (k & 1) ? a.d0 : a.d1;
            auto dirv = [&](long cc) -> double {
                const double z = a.r[cc] / va.diag[cc];
                return k > 0 ? z + beta * dold[cc] : z;
            };
            double dq[1] = {0.0};
            FOR_CELLS {
                const long c = ix(i, j, pc);
                if (!fluid(i, j)) continue;
                const int iE = (periodic && i == nx) ? 1 : i + 1;
                const int iW = (periodic && i == 1) ? nx : i - 1;
                const double ce = va.cE[c];
                const double cw = (periodic && i == 1) ? va.cE[ix(nx, j, pc)] : cW(i, j);
                const double cn = va.cN[c], csv = j > 1 ? va.cN[c - 1] : 0.0;
                const double dc = dirv(c);
                double q = va.diag[c] * dc;
                if (ce != 0.0) q -= ce * dirv(ix(iE, j, pc));
                if (cw != 0.0) q -= cw * dirv(ix(iW, j, pc));
                if (cn != 0.0) q -= cn * dirv(c + 1);
                if (csv != 0.0) q -= csv * dirv(c - 1);
                dnew[c] = dc;
                a.q[c] = q;
                dq[0] += dc * q;
            }
            grid_sum<1>(dq, gs, lds, epoch);
            const double alpha = rz / dq[0];
            double s2[2] = {0.0, 0.0};
            FOR_CELLS {
                const long c = ix(i, j, pc);
                if (!fluid(i, j)) continue;
                a.x[c] += alpha * dnew[c];
                const double rv = a.r[c] - alpha * a.q[c];
                a.r[c] = rv;
                s2[0] += rv * rv;
                s2[1] += rv * (rv / va.diag[c]);
            }
            grid_sum<2>(s2, gs, lds, epoch);
            rr = s2[0];
            rz_prev = rz;
            rz = s2[1];
            ++iters;
        }
        if (!conv) info = iters > 0 ? iters : a.maxiter;
    }
    if (lead) {
        st->cg_iters = iters;
        st->cg_info = info;
        st->cg_relres = s3[0] > 0.0 ? sqrt(rr) / sqrt(s3[0]) : 0.0;
    }

    // ---- p with the component means removed, ghost ring ----
    if (a.ncomp > 0) {
#pragma unroll
        for (int k = 0; k < CFD_MAX_COMP; ++k) cs[k] = 0.0;
        FOR_CELLS {
            const long c = ix(i, j, pc);
            const int cm = a.comp[c];
            const double xv = a.x[c];
#pragma unroll
            for (int k = 0; k < CFD_MAX_COMP; ++k) cs[k] += (cm == k + 1) ? xv : 0.0;
        }
        grid_sum<CFD_MAX_COMP>(cs, gs, lds, epoch);
#pragma unroll
        for (int k = 0; k < CFD_MAX_COMP; ++k)
            mean[k] = k < a.ncomp ? cs[k] / a.comp_count[k] : 0.0;
    } else {
#pragma unroll
        for (int k = 0; k < CFD_MAX_COMP; ++k) mean[k] = 0.0;
    }
    auto pval = [&](int i, int j) -> double {
        const long c = ix(i, j, pc);
        if (!(a.code[c] & 63)) return 0.0;
        const int cm = a.comp[c];
        const double xv = a.x[c];
        return cm ? xv - sel_comp(mean, cm) : xv;
    };
    for (int i = i0; i < i1; ++i)
        for (int j = tid; j <= ny + 1; j += BS)
            a.p[ix(i, j, pc)] = (j >= 1 && j <= ny) ? pval(i, j) : 0.0;
    if (blockIdx.x == 0)
        for (int j = tid; j <= ny + 1; j += BS)
            a.p[ix(0, j, pc)] = (periodic && j >= 1 && j <= ny) ? pval(nx, j) : 0.0;
    if (blockIdx.x == gridDim.x - 1)
        for (int j = tid; j <= ny + 1; j += BS)
            a.p[ix(nx + 1, j, pc)] = (periodic && j >= 1 && j <= ny) ? pval(1, j) : 0.0;
    grid_sync(gs, lds, epoch);

    // ---- u = F - dt (1/rho_f) grad p, residual monitor (solver.py:283-323) ----
    const double* p = a.p;
    double res[1] = {0.0};
    for (int i = i0; i < i1; ++i) {
        for (int j = 1 + tid; j <= ny; j += BS) {
            const long cu = ix(i, j, g.pu);
            const double F = a.F[cu];
            double un;
            bool uf = false;
            if (i < nx) {
                uf = fluid(i, j) && fluid(i + 1, j);
                un = uf ? F - dt * bx_face(va.rho, g, i, j) * (p[ix(i + 1, j, pc)] - p[ix(i, j, pc)]) / g.dx
                        : F;
            } else if (periodic) {
                uf = fluid(nx, j) && fluid(1, j);
                un = uf ? F - dt * bx_face(va.rho, g, nx, j) * (p[ix(1, j, pc)] - p[ix(nx, j, pc)]) / g.dx
                        : F;
            } else {
                un = (a.code[ix(nx, j, pc)] & 64)
                         ? F + 2 * dt * bx_face(va.rho, g, nx, j) * p[ix(nx, j, pc)] / g.dx
                         : F;
            }
            if (uf) res[0] = nanmax(res[0], fabs(un - a.u[cu]));
            a.u[cu] = un;
            if (periodic && i == nx) a.u[ix(0, j, g.pu)] = un;
        }
        for (int j = tid; j <= ny; j += BS) {
            const long cv = ix(i, j, g.pv);
            const double G = a.G[cv];
            double vn;
            bool vf = false;
            if (j >= 1 && j < ny) {
                vf = fluid(i, j) && fluid(i, j + 1);
                vn = vf ? G - dt * by_face(va.rho, g, i, j) * (p[ix(i, j + 1, pc)] - p[ix(i, j, pc)]) / g.dy
                        : G;
            } else if (periodic) {
                continue;
            } else if (j == 0) {
                vn = (a.code[ix(i, 1, pc)] & 128)
                         ? G - 2 * dt * by_face(va.rho, g, i, 0) * p[ix(i, 1, pc)] / g.dy
                         : G;
            } else {
                vn = (a.code[ix(i, ny, pc)] & 128)
                         ? G + 2 * dt * by_face(va.rho, g, i, ny) * p[ix(i, ny, pc)] / g.dy
                         : G;
            }
            if (vf) res[0] = nanmax(res[0], fabs(vn - a.v[cv]));
            a.v[cv] = vn;
        }
    }
    if (blockIdx.x == 0 && !periodic) {
        for (int j = 1 + tid; j <= ny; j += BS) {
            const long cu = ix(0, j, g.pu);
            const double F = a.F[cu];
            a.u[cu] = (a.code[ix(1, j, pc)] & 64)
                          ? F - 2 * dt * bx_face(va.rho, g, 0, j) * p[ix(1, j, pc)] / g.dx
                          : F;
        }
    }
    grid_max<1>(res, gs, lds, epoch);
    if (lead && a.dt_override <= 0.0) st->residual = res[0];
    if (periodic && blockIdx.x == gridDim.x - 1 && tid < 2) {
        const int j = tid == 0 ? 0 : ny + 1;
        a.u[ix(0, j, g.pu)] = a.u[ix(nx, j, g.pu)];
    }
    if (a.zero_p_after) {
        for (int i = i0; i < i1; ++i)
            for (int j = tid; j <= ny + 1; j += BS) a.p[ix(i, j, pc)] = 0.0;
        if (blockIdx.x == 0)
            for (int j = tid; j <= ny + 1; j += BS) a.p[ix(0, j, pc)] = 0.0;
        if (blockIdx.x == gridDim.x - 1)
            for (int j = tid; j <= ny + 1; j += BS) a.p[ix(nx + 1, j, pc)] = 0.0;
    }
#undef FOR_CELLS
}
