// Explicit stencils of the projection step on gfx950: domain and obstacle
// boundary conditions, the time-step and gamma control, the fused momentum
// predictor (momentum tendencies + Boussinesq source + AB2 + F/G write +
// outflow extrapolation), TVD scalar transport with ghost fill, and the
// per-step monitors.
//
// Every floating-point expression keeps the operand order of the reference
// (cfd2d/kernels.py, solver.py, boundary.py, models.py) and the library is
// compiled with -ffp-contract=off, so these kernels reproduce the reference's
// IEEE double results bit for bit.
#pragma clang fp contract(off)
#include "cfd_kernels.h"

// ---------------------------------------------------------------------------
// Momentum tendencies at one face (kernels.py:84-155).  VIS selects the
// diffusion term: 0 the constant viscosity nu0 (Laplacian form, lines
// 104-105); 1 the stress form div(nu (grad u + grad u^T)) of the cell field
// nuf (lines 96-103, 131-138); 2 the stress form of nu0 + nuf (nu_eff =
// nu + nu_t, solver.py:332-333, formed where it is read).
// ---------------------------------------------------------------------------
template <int VIS>
__device__ __forceinline__ double nu_at(const double* __restrict__ nuf, double nu0, int pc, int i,
                                        int j) {
    if (VIS == 1) return nuf[ix(i, j, pc)];
    return nu0 + nuf[ix(i, j, pc)];
}

template <int VIS>
__device__ __forceinline__ double mom_u_face(const double* __restrict__ u,
                                             const double* __restrict__ v, int pu, int pv,
                                             int i, int ip, int j, double nu0,
                                             const double* __restrict__ nuf, int pc, double gamma,
                                             double idx2, double idy2, double idx, double idy,
                                             const double* __restrict__ rho = nullptr) {
    const double uc = u[ix(i, j, pu)], ue = u[ix(ip, j, pu)], uw = u[ix(i - 1, j, pu)];
    const double un = u[ix(i, j + 1, pu)], us = u[ix(i, j - 1, pu)];
    double diff;
    if (VIS == 0) {
        diff = nu0 * ((ue - 2.0 * uc + uw) * idx2 + (un - 2.0 * uc + us) * idy2);
    } else {
        const double nc = nu_at<VIS>(nuf, nu0, pc, i, j), ne = nu_at<VIS>(nuf, nu0, pc, i + 1, j);
        const double nn = nu_at<VIS>(nuf, nu0, pc, i, j + 1);
        const double nne = nu_at<VIS>(nuf, nu0, pc, i + 1, j + 1);
        const double ns = nu_at<VIS>(nuf, nu0, pc, i, j - 1);
        const double nse = nu_at<VIS>(nuf, nu0, pc, i + 1, j - 1);
        const double nu_n = 0.25 * (nc + ne + nn + nne);
        const double nu_s = 0.25 * (nc + ne + ns + nse);
        diff = (2.0 * ne * (ue - uc) - 2.0 * nc * (uc - uw)) * idx2 +
               idy * (nu_n * ((un - uc) * idy + (v[ix(i + 1, j, pv)] - v[ix(i, j, pv)]) * idx) -
                      nu_s * ((uc - us) * idy +
                              (v[ix(i + 1, j - 1, pv)] - v[ix(i, j - 1, pv)]) * idx));
        // variable density: nuf is the dynamic viscosity (kernels.py:104-105)
        if (rho) diff /= 0.5 * (rho[ix(i, j, pc)] + rho[ix(i + 1, j, pc)]);
    }
    const double a_e = 0.5 * (uc + ue), a_w = 0.5 * (uw + uc);
    const double du2dx = idx * (a_e * a_e - a_w * a_w)
        + gamma * idx * (fabs(a_e) * 0.5 * (uc - ue) - fabs(a_w) * 0.5 * (uw - uc));
    const double v_n = 0.5 * (v[ix(i, j, pv)] + v[ix(i + 1, j, pv)]);
    const double v_s = 0.5 * (v[ix(i, j - 1, pv)] + v[ix(i + 1, j - 1, pv)]);
    const double duvdy = idy * (v_n * 0.5 * (uc + un) - v_s * 0.5 * (us + uc))
        + gamma * idy * (fabs(v_n) * 0.5 * (uc - un) - fabs(v_s) * 0.5 * (us - uc));
    return diff - du2dx - duvdy;
}

template <int VIS>
__device__ __forceinline__ double mom_v_face(const double* __restrict__ u,
                                             const double* __restrict__ v, int pu, int pv,
                                             int i, int j, double nu0,
                                             const double* __restrict__ nuf, int pc, double gamma,
                                             double idx2, double idy2, double idx, double idy,
                                             const double* __restrict__ rho = nullptr) {
    const double vc = v[ix(i, j, pv)], ve = v[ix(i + 1, j, pv)], vw = v[ix(i - 1, j, pv)];
    const double vn = v[ix(i, j + 1, pv)], vs = v[ix(i, j - 1, pv)];
    double diff;
    if (VIS == 0) {
        diff = nu0 * ((ve - 2.0 * vc + vw) * idx2 + (vn - 2.0 * vc + vs) * idy2);
    } else {
        const double nc = nu_at<VIS>(nuf, nu0, pc, i, j), nn = nu_at<VIS>(nuf, nu0, pc, i, j + 1);
        const double ne = nu_at<VIS>(nuf, nu0, pc, i + 1, j);
        const double nne = nu_at<VIS>(nuf, nu0, pc, i + 1, j + 1);
        const double nw = nu_at<VIS>(nuf, nu0, pc, i - 1, j);
        const double nnw = nu_at<VIS>(nuf, nu0, pc, i - 1, j + 1);
        const double nu_e = 0.25 * (nc + ne + nn + nne);
        const double nu_w = 0.25 * (nc + nw + nn + nnw);
        diff = (2.0 * nn * (vn - vc) - 2.0 * nc * (vc - vs)) * idy2 +
               idx * (nu_e * ((ve - vc) * idx + (u[ix(i, j + 1, pu)] - u[ix(i, j, pu)]) * idy) -
                      nu_w * ((vc - vw) * idx +
                              (u[ix(i - 1, j + 1, pu)] - u[ix(i - 1, j, pu)]) * idy));
        if (rho) diff /= 0.5 * (rho[ix(i, j, pc)] + rho[ix(i, j + 1, pc)]);   // lines 139-140
    }
    const double b_n = 0.5 * (vc + vn), b_s = 0.5 * (vs + vc);
    const double dv2dy = idy * (b_n * b_n - b_s * b_s)
        + gamma * idy * (fabs(b_n) * 0.5 * (vc - vn) - fabs(b_s) * 0.5 * (vs - vc));
    const double u_e = 0.5 * (u[ix(i, j, pu)] + u[ix(i, j + 1, pu)]);
    const double u_w = 0.5 * (u[ix(i - 1, j, pu)] + u[ix(i - 1, j + 1, pu)]);
    const double duvdx = idx * (u_e * 0.5 * (vc + ve) - u_w * 0.5 * (vw + vc))
        + gamma * idx * (fabs(u_e) * 0.5 * (vc - ve) - fabs(u_w) * 0.5 * (vw - vc));
    return diff - duvdx - dv2dy;
}

// Standalone momentum_tendencies (kernels.py:44-155): one thread per entry of
// the (nx+2) x (ny+2) index box; writes every entry of hu and hv (zero off the
// fluid faces, as kernels.py:74-79 does).
// VIS 0: constant nu0; VIS 1: the stress form of the cell field nuf (nu_eff).
template <int VIS>
__global__ void k_momentum_tendencies(const double* __restrict__ u, const double* __restrict__ v,
                                      const uint8_t* __restrict__ fl, Geo g, double nu0,
                                      const double* __restrict__ nuf, double gamma,
                                      double* __restrict__ hu, double* __restrict__ hv) {
    const int2 rc_ = xcd_rowgrid();
    const int j = rc_.x * blockDim.x + threadIdx.x;
    const int i = rc_.y;
    if (j > g.ny + 1) return;
    const double idx2 = 1.0 / (g.dx * g.dx), idy2 = 1.0 / (g.dy * g.dy);
    const double idx = 1.0 / g.dx, idy = 1.0 / g.dy;
    const int nx = g.nx, ny = g.ny;
    if (i <= nx) {
        const int imax = g.periodic ? nx : nx - 1;
        double h = 0.0;
        if (i >= 1 && i <= imax && j >= 1 && j <= ny && fl[ix(i, j, g.pc)] & 1 &&
            fl[ix(i + 1, j, g.pc)] & 1) {
            const int ip = i < nx ? i + 1 : 1;
            h = mom_u_face<VIS>(u, v, g.pu, g.pv, i, ip, j, nu0, nuf, g.pc, gamma, idx2, idy2,
                                idx, idy);
        }
        hu[ix(i, j, g.pu)] = h;
    }
    if (j <= ny) {
        double h = 0.0;
        if (i >= 1 && i <= nx && j >= 1 && j < ny && fl[ix(i, j, g.pc)] & 1 &&
            fl[ix(i, j + 1, g.pc)] & 1)
            h = mom_v_face<VIS>(u, v, g.pu, g.pv, i, j, nu0, nuf, g.pc, gamma, idx2, idy2, idx,
                                idy);
        hv[ix(i, j, g.pv)] = h;
    }
}

// ---------------------------------------------------------------------------
// Fused predictor (solver.py:338-368, models.py:386-396, boundary.py:225-242)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_predictor(PredArgs a, Geo g) {
    const int2 rc_ = xcd_rowgrid();
    const int j = rc_.x * blockDim.x + threadIdx.x;
    const int i = rc_.y;
    const int nx = g.nx, ny = g.ny;
    if (j > ny + 1) return;
    const DevState* st = a.st;
    if (st->done) return;
    const double dt = st->dt, gamma = st->gamma;
    const bool ab2 = a.scheme_ab2 && st->has_old;
    const double A = st->ab_a, B = st->ab_b;
    const double idx2 = 1.0 / (g.dx * g.dx), idy2 = 1.0 / (g.dy * g.dy);
    const double idx = 1.0 / g.dx, idy = 1.0 / g.dy;
    const double fx = a.fx + st->flow_force;
    const double fy = a.fy;
    const uint8_t* fl = a.fl;

    // ---- u face (i, j), i in 0..nx ----
    if (i <= nx) {
        // u_fluid (grid.py:138-142 with solver.py:138-139): row 0 is never a fluid face
        const bool uf = i >= 1 && (fl[ix(i, j, g.pc)] & 1) && (fl[ix(i + 1, j, g.pc)] & 1);
        double h = 0.0;
        if (uf) {   // implies 1 <= i <= (periodic ? nx : nx-1) and 1 <= j <= ny
            const int ip = i < nx ? i + 1 : 1;
            if (a.rho)
                h = mom_u_face<1>(a.u, a.v, g.pu, g.pv, i, ip, j, 0.0, a.nut, g.pc, gamma, idx2,
                                  idy2, idx, idy, a.rho);
            else
                h = a.nut ? mom_u_face<2>(a.u, a.v, g.pu, g.pv, i, ip, j, a.nu0, a.nut, g.pc,
                                          gamma, idx2, idy2, idx, idy)
                          : mom_u_face<0>(a.u, a.v, g.pu, g.pv, i, ip, j, a.nu0, a.nut, g.pc,
                                          gamma, idx2, idy2, idx, idy);
        }
        if (uf && a.kappa) {   // CSF surface tension over the face density (vof.py:235-240)
            const int c0 = ix(i, j, g.pc), c1 = ix(i + 1, j, g.pc);
            const double fxs = a.sigma * 0.5 * (a.kappa[c0] + a.kappa[c1]) *
                               (a.alpha[c1] - a.alpha[c0]) / g.dx;
            h = h + fxs / (0.5 * (a.rho[c0] + a.rho[c1]));
        }
        if (uf && a.T && a.gx != 0.0) {
            const double Tf = 0.5 * (a.T[ix(i, j, g.pc)] + a.T[ix(i + 1, j, g.pc)]) - a.Tref;
            h += -a.gx * Tf;
        }
        const double uval = a.u[ix(i, j, g.pu)];
        double F = uval;
        if (uf) {
            const double H = ab2 ? A * h + B * a.hu_old[ix(i, j, g.pu)] : h;
            F = uval + dt * H;
            if (fx != 0.0) F = F + dt * fx;
        }
        a.hu[ix(i, j, g.pu)] = h;
        // outflow / periodic extrapolation of the predictor (boundary.py:225-242)
        bool skip = false;
        if (i == 0 && g.periodic) skip = true;
        if (i == 0 && a.ext[0] && j >= 1 && j <= ny && a.open[0][j - 1]) skip = true;
        if (i == nx && a.ext[1] && j >= 1 && j <= ny && a.open[1][j - 1]) skip = true;
        if (!skip) a.F[ix(i, j, g.pu)] = F;
        if (i == nx && g.periodic) a.F[ix(0, j, g.pu)] = F;
        if (i == 1 && a.ext[0] && j >= 1 && j <= ny && a.open[0][j - 1]) a.F[ix(0, j, g.pu)] = F;
        if (i == nx - 1 && a.ext[1] && j >= 1 && j <= ny && a.open[1][j - 1])
            a.F[ix(nx, j, g.pu)] = F;
    }
    // ---- v face (i, j), j in 0..ny, i in 0..nx+1 ----
    if (j <= ny) {
        const bool vf = (fl[ix(i, j, g.pc)] & 1) && (fl[ix(i, j + 1, g.pc)] & 1);
        double h = 0.0;
        if (vf && i >= 1 && i <= nx && j >= 1 && j < ny)
            h = a.rho ? mom_v_face<1>(a.u, a.v, g.pu, g.pv, i, j, 0.0, a.nut, g.pc, gamma, idx2,
                                      idy2, idx, idy, a.rho)
                : a.nut ? mom_v_face<2>(a.u, a.v, g.pu, g.pv, i, j, a.nu0, a.nut, g.pc, gamma,
                                        idx2, idy2, idx, idy)
                        : mom_v_face<0>(a.u, a.v, g.pu, g.pv, i, j, a.nu0, a.nut, g.pc, gamma,
                                        idx2, idy2, idx, idy);
        if (vf && a.kappa) {
            const int c0 = ix(i, j, g.pc), c1 = ix(i, j + 1, g.pc);
            const double fys = a.sigma * 0.5 * (a.kappa[c0] + a.kappa[c1]) *
                               (a.alpha[c1] - a.alpha[c0]) / g.dy;
            h = h + fys / (0.5 * (a.rho[c0] + a.rho[c1]));
        }
        if (vf && a.T && a.gy != 0.0) {
            const double Tf = 0.5 * (a.T[ix(i, j, g.pc)] + a.T[ix(i, j + 1, g.pc)]) - a.Tref;
            h += -a.gy * Tf;
        }
        const double vval = a.v[ix(i, j, g.pv)];
        double G = vval;
        if (vf) {
            const double H = ab2 ? A * h + B * a.hv_old[ix(i, j, g.pv)] : h;
            G = vval + dt * H;
            if (fy != 0.0) G = G + dt * fy;
        }
        a.hv[ix(i, j, g.pv)] = h;
        const bool inner_i = i >= 1 && i <= nx;
        bool skip = false;
        if (j == 0 && a.ext[2] && inner_i && a.open[2][i - 1]) skip = true;
        if (j == ny && a.ext[3] && inner_i && a.open[3][i - 1]) skip = true;
        if (!skip) a.G[ix(i, j, g.pv)] = G;
        if (j == 1 && a.ext[2] && inner_i && a.open[2][i - 1]) a.G[ix(i, 0, g.pv)] = G;
        if (j == ny - 1 && a.ext[3] && inner_i && a.open[3][i - 1]) a.G[ix(i, ny, g.pv)] = G;
    }
}

// Copy F = u, G = v with the outflow / periodic extrapolation
// (solver.py:199-201, the predictor of the initial projection).
__global__ void k_copy_predictor(PredArgs a, Geo g) {
    const int2 rc_ = xcd_rowgrid();
    const int j = rc_.x * blockDim.x + threadIdx.x;
    const int i = rc_.y;
    const int nx = g.nx, ny = g.ny;
    if (j > ny + 1) return;
    if (i <= nx) {
        const double F = a.u[ix(i, j, g.pu)];
        bool skip = (i == 0 && g.periodic) ||
                    (i == 0 && a.ext[0] && j >= 1 && j <= ny && a.open[0][j - 1]) ||
                    (i == nx && a.ext[1] && j >= 1 && j <= ny && a.open[1][j - 1]);
        if (!skip) a.F[ix(i, j, g.pu)] = F;
        if (i == nx && g.periodic) a.F[ix(0, j, g.pu)] = F;
        if (i == 1 && a.ext[0] && j >= 1 && j <= ny && a.open[0][j - 1]) a.F[ix(0, j, g.pu)] = F;
        if (i == nx - 1 && a.ext[1] && j >= 1 && j <= ny && a.open[1][j - 1])
            a.F[ix(nx, j, g.pu)] = F;
    }
    if (j <= ny) {
        const double G = a.v[ix(i, j, g.pv)];
        const bool inner_i = i >= 1 && i <= nx;
        bool skip = (j == 0 && a.ext[2] && inner_i && a.open[2][i - 1]) ||
                    (j == ny && a.ext[3] && inner_i && a.open[3][i - 1]);
        if (!skip) a.G[ix(i, j, g.pv)] = G;
        if (j == 1 && a.ext[2] && inner_i && a.open[2][i - 1]) a.G[ix(i, 0, g.pv)] = G;
        if (j == ny - 1 && a.ext[3] && inner_i && a.open[3][i - 1]) a.G[ix(i, ny, g.pv)] = G;
    }
}

// ---------------------------------------------------------------------------
// Domain boundary conditions (boundary.py:143-223); the phases follow the
// reference's statement order where they touch, so ghost corners come out
// identical.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double bc_normal(int kind, double cur, const double* inflow,
                                            const uint8_t* open, int k) {
    switch (kind) {
    case BC_NOSLIP:
    case BC_SLIP: return 0.0;
    case BC_INFLOW: return inflow[k];
    case BC_OUTFLOW: return open[k] ? cur : 0.0;
    default: return cur;
    }
}
__device__ __forceinline__ double bc_tangential(int kind, double inner, double wall, double cur) {
    switch (kind) {
    case BC_NOSLIP: return 2.0 * wall - inner;
    case BC_SLIP: return inner;
    case BC_INFLOW: return -inner;
    case BC_OUTFLOW: return inner;
    default: return cur;
    }
}

// indices first..last over the work-group, R per thread at a time: the R
// entries' loads all come before any of their stores, which the compiler could
// not reorder itself (the stores might alias the later loads); no index of a
// phase reads what another index of it writes
template <int R, typename L, typename S>
__device__ __forceinline__ void bc_batch(int first, int last, L load, S store) {
    using V = decltype(load(first));
    const int T = blockDim.x;
    for (int k0 = first + (int)threadIdx.x; k0 <= last; k0 += R * T) {
        V val[R];
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (k0 + r * T <= last) val[r] = load(k0 + r * T);
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (k0 + r * T <= last) store(k0 + r * T, val[r]);
    }
}

// the normal component's raw inputs at boundary entry k of side s
struct BCNormalIn {
    double cur, in;
    int open;
};
__device__ __forceinline__ BCNormalIn bc_normal_in(const BCArgs& a, int s, double cur, int k) {
    BCNormalIn e;
    e.cur = cur;
    e.in = a.kind[s] == BC_INFLOW ? a.inflow[s][k] : 0.0;
    e.open = a.kind[s] == BC_OUTFLOW ? a.open[s][k] : 1;
    return e;
}
__device__ __forceinline__ double bc_normal_of(int kind, const BCNormalIn& e) {
    switch (kind) {
    case BC_NOSLIP:
    case BC_SLIP: return 0.0;
    case BC_INFLOW: return e.in;
    case BC_OUTFLOW: return e.open ? e.cur : 0.0;
    default: return e.cur;
    }
}

// bottom / top of column i (the second statement group of boundary.py:143-223),
// then its faces closed by an obstacle behind the boundary (boundary.py:219-223)
struct BT {
    BCNormalIn s, n;
    double u0, u1, un, un1;
    int lo, hi;
};
__device__ __forceinline__ BT bc_bt_load(const BCArgs& a, const Geo& g, int i) {
    const int ny = g.ny, pu = g.pu, pv = g.pv;
    BT r;
    const int k = i >= 1 ? i - 1 : 0;
    r.s = bc_normal_in(a, 2, a.v[ix(i, 0, pv)], k);
    r.n = bc_normal_in(a, 3, a.v[ix(i, ny, pv)], k);
    r.u0 = a.u[ix(i, 0, pu)];
    r.u1 = a.u[ix(i, 1, pu)];
    r.un = a.u[ix(i, ny, pu)];
    r.un1 = a.u[ix(i, ny + 1, pu)];
    r.lo = i >= 1 ? a.open[2][k] : 1;
    r.hi = i >= 1 ? a.open[3][k] : 1;
    return r;
}
__device__ __forceinline__ void bc_bt_store(const BCArgs& a, const Geo& g, int i, const BT& r) {
    const int ny = g.ny, pu = g.pu, pv = g.pv;
    if (i >= 1) {
        a.v[ix(i, 0, pv)] = r.lo ? bc_normal_of(a.kind[2], r.s) : 0.0;
        a.v[ix(i, ny, pv)] = r.hi ? bc_normal_of(a.kind[3], r.n) : 0.0;
    }
    a.u[ix(i, 0, pu)] = bc_tangential(a.kind[2], r.u1, a.vel[2], r.u0);
    a.u[ix(i, ny + 1, pu)] = bc_tangential(a.kind[3], r.un, a.vel[3], r.un1);
}

// Work-group 0: the periodic copies and the left / right group over the rows
// i = 0, 1, nx, nx + 1, then the bottom / top group on the columns those rows
// share with it (i = 0, 1, nx), then the closed left / right faces.  The other
// work-groups: the bottom / top group on columns 2 .. nx - 1, BC_COLS each,
// one column per thread -- those columns neither read nor write anything the
// left / right group touches, so the result is that of the statement order.
__global__ void __launch_bounds__(256) k_bc_domain(BCArgs a, Geo g) {
    const int nx = g.nx, ny = g.ny, pu = g.pu, pv = g.pv;
    double* u = a.u;
    double* v = a.v;
    if (a.st && a.st->done) return;
    if (blockIdx.x > 0) {
        const int i = 2 + (int)(blockIdx.x - 1) * BC_COLS + (int)threadIdx.x;
        if (threadIdx.x < BC_COLS && i <= nx - 1) bc_bt_store(a, g, i, bc_bt_load(a, g, i));
        return;
    }
    const int t = threadIdx.x, T = blockDim.x;
    if (g.periodic) {
        for (int j = t; j <= ny + 1; j += T) u[ix(0, j, pu)] = u[ix(nx, j, pu)];
        for (int j = t; j <= ny; j += T) {
            v[ix(0, j, pv)] = v[ix(nx, j, pv)];
            v[ix(nx + 1, j, pv)] = v[ix(1, j, pv)];
        }
    }
    __syncthreads();
    // left / right
    struct LR {
        BCNormalIn w, e;
        double v0, v1, vn, vn1;
    };
    bc_batch<4>(
        0, ny,
        [&](int j) {
            LR r;
            const int k = j >= 1 ? j - 1 : 0;
            r.w = bc_normal_in(a, 0, u[ix(0, j, pu)], k);
            r.e = bc_normal_in(a, 1, u[ix(nx, j, pu)], k);
            r.v0 = v[ix(0, j, pv)];
            r.v1 = v[ix(1, j, pv)];
            r.vn = v[ix(nx, j, pv)];
            r.vn1 = v[ix(nx + 1, j, pv)];
            return r;
        },
        [&](int j, const LR& r) {
            if (j >= 1) {
                u[ix(0, j, pu)] = bc_normal_of(a.kind[0], r.w);
                u[ix(nx, j, pu)] = bc_normal_of(a.kind[1], r.e);
            }
            v[ix(0, j, pv)] = bc_tangential(a.kind[0], r.v1, a.vel[0], r.v0);
            v[ix(nx + 1, j, pv)] = bc_tangential(a.kind[1], r.vn, a.vel[1], r.vn1);
        });
    __syncthreads();
    // bottom / top on columns 0, 1 and nx (the others in the other work-groups)
    if (t < 3) {
        const int i = t == 0 ? 0 : (t == 1 ? 1 : nx);
        const bool dup = (t == 1 && nx < 1) || (t == 2 && nx <= 1);
        if (!dup) bc_bt_store(a, g, i, bc_bt_load(a, g, i));
    }
    __syncthreads();
    // faces closed by an obstacle behind the left / right boundary
    struct Open2 {
        int lo, hi;
    };
    if (!g.periodic) {
        bc_batch<4>(
            1, ny, [&](int j) { return Open2{a.open[0][j - 1], a.open[1][j - 1]}; },
            [&](int j, const Open2& o) {
                if (!o.lo) u[ix(0, j, pu)] = 0.0;
                if (!o.hi) u[ix(nx, j, pu)] = 0.0;
            });
    }
}

// ---------------------------------------------------------------------------
// Obstacle no-slip (kernels.py:158-209).  Each face reads only faces between
// two fluid cells, which this kernel never writes, so all faces update in
// parallel with the reference's result.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_obstacle_bc(double* __restrict__ u, double* __restrict__ v,
                              const uint8_t* __restrict__ fl, Geo g, const DevState* st) {
    const int2 rc_ = xcd_rowgrid();
    const int j = rc_.x * blockDim.x + threadIdx.x;
    const int i = rc_.y;
    const int nx = g.nx, ny = g.ny, pu = g.pu, pv = g.pv, pc = g.pc;
    if (j > ny + 1) return;
    if (st && st->done) return;
    const int imax = g.periodic ? nx : nx - 1;
    if (i >= 1 && i <= imax && j >= 1 && j <= ny) {
        const bool fL = fl[ix(i, j, pc)] & 1, fR = fl[ix(i + 1, j, pc)] & 1;
        if (!(fL && fR)) {
            double val;
            if (fL || fR) {
                val = 0.0;
            } else {
                double s = 0.0;
                int c = 0;
                if ((fl[ix(i, j + 1, pc)] & 1) && (fl[ix(i + 1, j + 1, pc)] & 1)) {
                    s += u[ix(i, j + 1, pu)];
                    ++c;
                }
                if ((fl[ix(i, j - 1, pc)] & 1) && (fl[ix(i + 1, j - 1, pc)] & 1)) {
                    s += u[ix(i, j - 1, pu)];
                    ++c;
                }
                val = c > 0 ? -s / c : 0.0;
            }
            u[ix(i, j, pu)] = val;
            if (g.periodic && i == nx) u[ix(0, j, pu)] = val;
        } else if (g.periodic && i == nx) {
            u[ix(0, j, pu)] = u[ix(nx, j, pu)];
        }
    }
    if (g.periodic && i == nx && (j == 0 || j == ny + 1)) u[ix(0, j, pu)] = u[ix(nx, j, pu)];
    if (i >= 1 && i <= nx && j >= 1 && j < ny) {
        const bool fB = fl[ix(i, j, pc)] & 1, fT = fl[ix(i, j + 1, pc)] & 1;
        if (!(fB && fT)) {
            double val;
            if (fB || fT) {
                val = 0.0;
            } else {
                double s = 0.0;
                int c = 0;
                if ((fl[ix(i + 1, j, pc)] & 1) && (fl[ix(i + 1, j + 1, pc)] & 1)) {
                    s += v[ix(i + 1, j, pv)];
                    ++c;
                }
                if ((fl[ix(i - 1, j, pc)] & 1) && (fl[ix(i - 1, j + 1, pc)] & 1)) {
                    s += v[ix(i - 1, j, pv)];
                    ++c;
                }
                val = c > 0 ? -s / c : 0.0;
            }
            v[ix(i, j, pv)] = val;
        }
    }
}

// ---------------------------------------------------------------------------
// Time step and gamma (solver.py:209-253): max |u|, max |v| over the whole
// arrays (ghosts included), then the last work-group computes dt, gamma and
// the AB2 weights into the device state.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long dbits(double x) {
    return (unsigned long long)__double_as_longlong(x);
}

// With turbulence (td.nut set) the viscous limit and the cell Reynolds number
// use nu_eff = nu + nu_t: its maximum over every cell and its minimum over
// the fluid cells (solver.py:216-218, 249-251; max(nu + x) = nu + max(x)
// exactly, rounding being monotonic), and k-omega SST adds the explicit
// diffusion limits of k and omega from their largest diffusivity on the
// fluid cells (models.py:204-208, turbulence.py:314-315).  Each block leaves
// max |u|, max |v|, max nu_t, -min nu_t (fluid), max D_k, max D_omega in
// part[6 b ..]; block 0 takes the maxima and sets the step.
__global__ void __launch_bounds__(RED_THREADS) k_dt_gamma(const double* __restrict__ u, const double* __restrict__ v, Geo g,
                           StepParams prm, TurbDt td, DevState* st, double* part,
                           unsigned* flags, SpinGuard spin) {
    __shared__ double lds[6 * (RED_THREADS / CFD_WAVE)];
    const int nx = g.nx, ny = g.ny;
    double m[2] = {0.0, 0.0};
    // turbulence: max nu_t, -min nu_t (fluid), max D_k, max D_omega (fluid)
    double q[4] = {0.0, -__builtin_inf(), 0.0, 0.0};
    if (!st->done) {
        // rows 0..nx+1 x columns 0..ny+1 (u has rows 0..nx, v columns 0..ny)
        const int w = ny + 2, n = (nx + 2) * w;
        struct UV {
            double au, av;
            bool iu, jv;
        };
        flat_reduce<8>(
            n, w, 0, 0,
            [&](int i, int j) {
                UV e;
                e.au = fabs(u[ix(i <= nx ? i : nx, j, g.pu)]);
                e.av = fabs(v[ix(i, j <= ny ? j : ny, g.pv)]);
                e.iu = i <= nx;
                e.jv = j <= ny;
                return e;
            },
            [&](const UV& e, bool ok) {
                m[0] = ok && e.iu ? nanmax(m[0], e.au) : m[0];
                m[1] = ok && e.jv ? nanmax(m[1], e.av) : m[1];
            });
        if (td.nut) {
            struct Turb {
                double nt, dk, dw;
                bool f;
            };
            flat_reduce<4>(
                n, w, 0, 0,
                [&](int i, int j) {
                    const long c = ix(i, j, g.pc);
                    Turb e;
                    e.f = td.fl[c] & 1;
                    // variable density: nu = mu / rho over the fluid cells
                    e.nt = td.rho ? td.nut[c] / td.rho[c] : td.nut[c];
                    e.dk = td.dk ? td.dk[c] : 0.0;
                    e.dw = td.dk ? td.dw[c] : 0.0;
                    return e;
                },
                [&](const Turb& e, bool ok) {
                    const bool f = ok && e.f;
                    q[0] = (f || (ok && !td.rho)) ? nanmax(q[0], e.nt) : q[0];
                    q[1] = f ? nanmax(q[1], -e.nt) : q[1];
                    if (td.dk) {
                        q[2] = f ? nanmax(q[2], e.dk) : q[2];
                        q[3] = f ? nanmax(q[3], e.dw) : q[3];
                    }
                });
        }
    }
    block_max<2>(m, lds);
    if (td.nut) {
        __syncthreads();
        block_max<4>(q, lds);
    }
    if (threadIdx.x == 0) {
        double* my = part + 6 * blockIdx.x;
        my[0] = m[0];
        my[1] = m[1];
#pragma unroll
        for (int k = 0; k < 4; ++k) my[2 + k] = q[k];
    }
    if (!gather_block0(flags, spin)) return;
    // block 0: the maxima over the blocks
    double t[6] = {0.0, 0.0, 0.0, -__builtin_inf(), 0.0, 0.0};
    for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x)
#pragma unroll
        for (int k = 0; k < 6; ++k) t[k] = nanmax(t[k], part[6 * b + k]);
    block_max<6>(t, lds);
    {
        if (threadIdx.x == 0) {
            const double umax_raw = t[0], vmax_raw = t[1];
            double nu_max = prm.nu, nu_ref = prm.nu, lim_turb = __builtin_inf();
            if (td.nut) {
                const double nt_max = t[2], nt_min = -t[3], dk = t[4], dw = t[5];
                nu_max = td.rho ? nt_max : prm.nu + nt_max;
                nu_ref = td.rho ? nt_min : prm.nu + nt_min;
                if (!st->done) st->nut_max = nt_max;
                if (td.dk) {
                    const double h2 = 1.0 / (g.dx * g.dx) + 1.0 / (g.dy * g.dy);
                    const double lk = dk > 0.0 ? 0.5 / (dk * h2) : __builtin_inf();
                    const double lw = dw > 0.0 ? 0.5 / (dw * h2) : __builtin_inf();
                    lim_turb = fmin(lk, lw);
                }
            }
            if (!st->done) {
                double dt;
                if (prm.dt_fixed > 0.0) {
                    dt = prm.dt_fixed;
                } else {
                    const double umax = fmax(umax_raw, 1e-12), vmax = fmax(vmax_raw, 1e-12);
                    double dt_visc = 0.5 / (nu_max * (1.0 / (g.dx * g.dx) + 1.0 / (g.dy * g.dy)));
                    double dt_conv = fmin(g.dx / umax, g.dy / vmax);
                    if (prm.scheme_ab2) {
                        dt_visc *= 0.5;
                        dt_conv *= 0.8;
                    }
                    dt = prm.tau * fmin(dt_visc, dt_conv);
                    if (prm.dt_model < 1e300) dt = fmin(dt, prm.tau * prm.dt_model);
                    if (lim_turb < 1e300) dt = fmin(dt, prm.tau * lim_turb);
                    if (prm.clamp_t_end && st->t + dt > prm.t_end) dt = prm.t_end - st->t;
                }
                double gamma;
                if (prm.gamma_auto) {
                    if (!prm.scheme_ab2) {
                        const double cfl = fmax(umax_raw * dt / g.dx, vmax_raw * dt / g.dy);
                        gamma = fmin(1.0, fmax(0.05, 1.2 * cfl));
                    } else {
                        const double re_cell = fmax(umax_raw * g.dx, vmax_raw * g.dy) / nu_ref;
                        gamma = fmin(1.0, fmax(0.05, 0.015 * re_cell));
                    }
                } else {
                    gamma = prm.gamma_fixed;
                }
                st->umax = umax_raw;
                st->vmax = vmax_raw;
                st->dt = dt;
                st->gamma = gamma;
                if (prm.scheme_ab2 && st->has_old) {
                    const double r = dt / st->dt_old;
                    st->ab_a = 1.0 + 0.5 * r;
                    st->ab_b = -0.5 * r;
                } else {
                    st->ab_a = 1.0;
                    st->ab_b = 0.0;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// TVD scalar transport (kernels.py:215-391, models.py:186-208) in gather form:
// every fluid cell adds its four face fluxes in the reference's order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double limiter_psi(double r, int kind) {
    switch (kind) {
    case 0: return 0.0;
    case 1: return r <= 0.0 ? 0.0 : 2.0 * r / (1.0 + r);
    case 2: return r <= 0.0 ? 0.0 : fmin(r, 1.0);
    case 3: return r <= 0.0 ? 0.0 : fmax(fmin(2.0 * r, 1.0), fmin(r, 2.0));
    case 5: return r <= 0.0 ? 0.0 : fmin(2.0 * r, fmin(0.5 * (1.0 + r), 2.0));
    default: return 1.0;
    }
}

__device__ __forceinline__ double face_value(double phi_u, double phi_c, double phi_d,
                                             bool has_u, int kind) {
    if (kind == 0) return phi_c;
    if (kind == 4) return 0.5 * (phi_c + phi_d);
    if (!has_u) return phi_c;
    const double den = phi_d - phi_c;
    if (fabs(den) < 1e-300) return phi_c;
    const double r = (phi_c - phi_u) / den;
    return phi_c + 0.5 * limiter_psi(r, kind) * den;
}

template <bool CONST_D>
__device__ __forceinline__ double dcell(const ScalarArgs& a, int i, int j, int pc) {
    return CONST_D ? a.D0 : a.diff[ix(i, j, pc)];
}

// Signed contribution of x-face f (between cells f and f+1, f = nx wrapping to
// cell 1 when periodic) to cell `cell` (which is fluid).
template <bool CONST_D>
__device__ __forceinline__ double xface_term(const ScalarArgs& a, const Geo& g, int f, int j,
                                             int cell) {
    const int nx = g.nx, pc = g.pc;
    const double* phi = a.phi_in;
    const uint8_t* fl = a.fl;
    if (f == 0 || (f == nx && !g.periodic)) {
        int t;
        double sgn, val;
        if (f == 0) {
            sgn = 1.0;
            t = a.bt[0][j - 1];
            val = a.bv[0][j - 1];
        } else {
            sgn = -1.0;
            t = a.bt[1][j - 1];
            val = a.bv[1][j - 1];
        }
        const int P = cell;
        const double un = sgn * a.u[ix(f, j, g.pu)];
        const double php = phi[ix(P, j, pc)];
        const double phb = (t == 1 && un > 0.0) ? val : php;
        double flux_in = un * phb;
        if (t == 1) flux_in += dcell<CONST_D>(a, P, j, pc) * (val - php) / (0.5 * g.dx);
        else if (t == 2) flux_in += val;
        return flux_in / g.dx;
    }
    const int L = f, R = f < nx ? f + 1 : 1;
    const bool fL = fl[ix(L, j, pc)] & 1, fR = fl[ix(R, j, pc)] & 1;
    if (fL && fR) {
        const double uf = a.u[ix(f, j, g.pu)];
        int C, D, Ui;
        if (uf >= 0.0) {
            C = L; D = R; Ui = L - 1;
            if (Ui < 1) Ui = g.periodic ? nx : -1;
        } else {
            C = R; D = L; Ui = R + 1;
            if (Ui > nx) Ui = g.periodic ? 1 : -1;
        }
        const bool has_u = Ui >= 1 && (fl[ix(Ui, j, pc)] & 1);
        const double phu = has_u ? phi[ix(Ui, j, pc)] : 0.0;
        const double phf = face_value(phu, phi[ix(C, j, pc)], phi[ix(D, j, pc)], has_u, a.limiter);
        const double dl = dcell<CONST_D>(a, L, j, pc), dr = dcell<CONST_D>(a, R, j, pc);
        const double df = dl + dr > 0.0 ? 2.0 * dl * dr / (dl + dr) : 0.0;
        const double F = uf * phf - df * (phi[ix(R, j, pc)] - phi[ix(L, j, pc)]) / g.dx;
        return cell == L ? -(F / g.dx) : F / g.dx;
    }
    if (a.obs_t == 1) {
        const int P = cell;
        return dcell<CONST_D>(a, P, j, pc) * (a.obs_v - phi[ix(P, j, pc)]) / (0.5 * g.dx) / g.dx;
    }
    return 0.0;
}

template <bool CONST_D>
__device__ __forceinline__ double yface_term(const ScalarArgs& a, const Geo& g, int i, int f,
                                             int cell) {
    const int ny = g.ny, pc = g.pc;
    const double* phi = a.phi_in;
    const uint8_t* fl = a.fl;
    if (f == 0 || f == ny) {
        int t;
        double sgn, val;
        if (f == 0) {
            sgn = 1.0;
            t = a.bt[2][i - 1];
            val = a.bv[2][i - 1];
        } else {
            sgn = -1.0;
            t = a.bt[3][i - 1];
            val = a.bv[3][i - 1];
        }
        const int P = cell;
        const double vn = sgn * a.v[ix(i, f, g.pv)];
        const double php = phi[ix(i, P, pc)];
        const double phb = (t == 1 && vn > 0.0) ? val : php;
        double flux_in = vn * phb;
        if (t == 1) flux_in += dcell<CONST_D>(a, i, P, pc) * (val - php) / (0.5 * g.dy);
        else if (t == 2) flux_in += val;
        return flux_in / g.dy;
    }
    const int B = f, T = f + 1;
    const bool fB = fl[ix(i, B, pc)] & 1, fT = fl[ix(i, T, pc)] & 1;
    if (fB && fT) {
        const double vf = a.v[ix(i, f, g.pv)];
        int C, D, Uj;
        if (vf >= 0.0) { C = B; D = T; Uj = B - 1; }
        else { C = T; D = B; Uj = T + 1; }
        const bool has_u = 1 <= Uj && Uj <= ny && (fl[ix(i, Uj, pc)] & 1);
        const double phu = has_u ? phi[ix(i, Uj, pc)] : 0.0;
        const double phf = face_value(phu, phi[ix(i, C, pc)], phi[ix(i, D, pc)], has_u, a.limiter);
        const double db = dcell<CONST_D>(a, i, B, pc), dt_ = dcell<CONST_D>(a, i, T, pc);
        const double df = db + dt_ > 0.0 ? 2.0 * db * dt_ / (db + dt_) : 0.0;
        const double F = vf * phf - df * (phi[ix(i, T, pc)] - phi[ix(i, B, pc)]) / g.dy;
        return cell == B ? -(F / g.dy) : F / g.dy;
    }
    if (a.obs_t == 1) {
        const int P = cell;
        return dcell<CONST_D>(a, i, P, pc) * (a.obs_v - phi[ix(i, P, pc)]) / (0.5 * g.dy) / g.dy;
    }
    return 0.0;
}

// The loads of one face between two interior cells L and R (rows along x or
// columns along y), with the upwind cells UL = L - 1 (for a velocity >= 0)
// and UR = R + 1 (< 0), -1 where there is none: all of them issued before any
// is used, so that a face costs one memory round trip instead of the chain
// flags -> velocity -> upwind cell.  Absent cells load the cell R instead.
struct FaceIn {
    double vel, pL, pR, pUL, pUR, dL, dR;
    int UL, UR;
    bool fL, fR, fUL, fUR;
};

template <bool CONST_D>
__device__ __forceinline__ FaceIn face_load(const ScalarArgs& a, double vel, long cL, long cR,
                                            long cUL, long cUR, int UL, int UR) {
    FaceIn q;
    q.vel = vel;
    q.UL = UL;
    q.UR = UR;
    const long kUL = UL >= 0 ? cUL : cR, kUR = UR >= 0 ? cUR : cR;
    q.fL = a.fl[cL] & 1;
    q.fR = a.fl[cR] & 1;
    q.fUL = a.fl[kUL] & 1;
    q.fUR = a.fl[kUR] & 1;
    q.pL = a.phi_in[cL];
    q.pR = a.phi_in[cR];
    q.pUL = a.phi_in[kUL];
    q.pUR = a.phi_in[kUR];
    q.dL = CONST_D ? a.D0 : a.diff[cL];
    q.dR = CONST_D ? a.D0 : a.diff[cR];
    return q;
}

// x-face f (1 <= f, not the closed right boundary): L = f, R = f + 1 (1 past
// nx when periodic), the upwind rows as xface_term picks them
template <bool CONST_D>
__device__ __forceinline__ FaceIn face_load_x(const ScalarArgs& a, const Geo& g, int f, int j) {
    const int nx = g.nx, pc = g.pc;
    const int L = f, R = f < nx ? f + 1 : 1;
    int UL = L - 1, UR = R + 1;
    if (UL < 1) UL = g.periodic ? nx : -1;
    if (UR > nx) UR = g.periodic ? 1 : -1;
    return face_load<CONST_D>(a, a.u[ix(f, j, g.pu)], ix(L, j, pc), ix(R, j, pc),
                              ix(UL >= 0 ? UL : R, j, pc), ix(UR >= 0 ? UR : R, j, pc), UL, UR);
}

// y-face f (1 <= f <= ny - 1): B = f, T = f + 1, the upwind columns as
// yface_term picks them
template <bool CONST_D>
__device__ __forceinline__ FaceIn face_load_y(const ScalarArgs& a, const Geo& g, int i, int f) {
    const int ny = g.ny, pc = g.pc;
    const int B = f, T = f + 1;
    const int UB = B - 1 >= 1 ? B - 1 : -1, UT = T + 1 <= ny ? T + 1 : -1;
    return face_load<CONST_D>(a, a.v[ix(i, f, g.pv)], ix(i, B, pc), ix(i, T, pc),
                              ix(i, UB >= 0 ? UB : T, pc), ix(i, UT >= 0 ? UT : T, pc), UB, UT);
}

// the signed contribution of a loaded face to the fluid cell on its L side
// (left) or R side, h the spacing: xface_term's / yface_term's arithmetic
template <bool CONST_D>
__device__ __forceinline__ double face_term(const ScalarArgs& a, const FaceIn& q, bool left,
                                            double h) {
    if (q.fL && q.fR) {
        const bool pos = q.vel >= 0.0;
        const int U = pos ? q.UL : q.UR;
        const bool has_u = U >= 1 && (pos ? q.fUL : q.fUR);
        const double phu = has_u ? (pos ? q.pUL : q.pUR) : 0.0;
        const double phf = face_value(phu, pos ? q.pL : q.pR, pos ? q.pR : q.pL, has_u, a.limiter);
        const double df = q.dL + q.dR > 0.0 ? 2.0 * q.dL * q.dR / (q.dL + q.dR) : 0.0;
        const double F = q.vel * phf - df * (q.pR - q.pL) / h;
        return left ? -(F / h) : F / h;
    }
    if (a.obs_t == 1) {
        const double dP = left ? q.dL : q.dR, pP = left ? q.pL : q.pR;
        return dP * (a.obs_v - pP) / (0.5 * h) / h;
    }
    return 0.0;
}

// MODE 0: write the transport right-hand side into a.out (kernels.py output,
//         zero off the fluid interior).
// MODE 1: phi_out = phi_in + dt * rhs on fluid cells, phi_in elsewhere.
// The faces between interior cells are loaded first (face_load), the
// boundary faces by xface_term / yface_term.
template <bool CONST_D, int MODE>
__global__ void __launch_bounds__(256) k_scalar_transport(ScalarArgs a, Geo g) {
    const int2 rc_ = xcd_rowgrid();
    const int j = rc_.x * blockDim.x + threadIdx.x;
    const int i = rc_.y;
    const int nx = g.nx, ny = g.ny, pc = g.pc;
    if (j > ny + 1) return;
    if (MODE == 1 && a.st->done) return;
    const bool inner = i >= 1 && i <= nx && j >= 1 && j <= ny;
    bool fluid = false;
    double rhs = 0.0;
    if (inner) {
        fluid = a.fl[ix(i, j, pc)] & 1;
        // west, east, south, north (the reference's order)
        const int fw = (g.periodic && i == 1) ? nx : i - 1;
        const bool bw = fw == 0, be = i == nx && !g.periodic, bs = j == 1, bn = j == ny;
        FaceIn q[4];
        if (!bw) q[0] = face_load_x<CONST_D>(a, g, fw, j);
        if (!be) q[1] = face_load_x<CONST_D>(a, g, i, j);
        if (!bs) q[2] = face_load_y<CONST_D>(a, g, i, j - 1);
        if (!bn) q[3] = face_load_y<CONST_D>(a, g, i, j);
        if (fluid) {
            double out = bw ? xface_term<CONST_D>(a, g, fw, j, i)
                            : face_term<CONST_D>(a, q[0], fw == i, g.dx);   // (fw == i: nx = 1)
            out += be ? xface_term<CONST_D>(a, g, i, j, i) : face_term<CONST_D>(a, q[1], true, g.dx);
            out += bs ? yface_term<CONST_D>(a, g, i, j - 1, j) : face_term<CONST_D>(a, q[2], false, g.dy);
            out += bn ? yface_term<CONST_D>(a, g, i, j, j) : face_term<CONST_D>(a, q[3], true, g.dy);
            rhs = out;
        }
    }
    if (MODE == 0) {
        a.out[ix(i, j, pc)] = rhs;
    } else {
        const double p = a.phi_in[ix(i, j, pc)];
        a.phi_out[ix(i, j, pc)] = fluid ? p + a.st->dt * rhs : p;
    }
}

// Ghost and solid cells of a scalar after its update (models.py:139-183).
// phi_in still holds the previous values: ghost cells built from a solid edge
// cell use its previous value, as the reference fills ghosts before solids.
__global__ void __launch_bounds__(256) k_fill_ghosts(ScalarArgs a, Geo g) {
    const int2 rc_ = xcd_rowgrid();
    const int j = rc_.x * blockDim.x + threadIdx.x;
    const int i = rc_.y;
    const int nx = g.nx, ny = g.ny, pc = g.pc;
    if (j > ny + 1) return;
    if (a.st && a.st->done) return;
    const uint8_t* fl = a.fl;
    const double* pin = a.phi_in;
    double* po = a.phi_out;
    auto cur = [&](int ci, int cj) -> double {   // value after the update, before solid fill
        return (fl[ix(ci, cj, pc)] & 1) ? po[ix(ci, cj, pc)] : pin[ix(ci, cj, pc)];
    };
    auto D = [&](int ci, int cj) -> double {
        const double d = a.diff ? a.diff[ix(ci, cj, pc)] : a.D0;
        return fmax(d, 1e-300);
    };
    auto ghost = [&](double inner, int code, double val, double h, double Dl) -> double {
        if (code == 1) return 2 * val - inner;
        if (code == 2) return inner + val * h / Dl;
        return inner;
    };
    const bool ghost_col = (i == 0 || i == nx + 1);
    const bool ghost_row = (j == 0 || j == ny + 1);
    if (ghost_col) {
        if (g.periodic) {
            const int src = (i == 0) ? nx : 1;
            // phi[0,:] = phi[nx,:] precedes the bottom/top ghost update
            po[ix(i, j, pc)] = ghost_row ? pin[ix(src, j, pc)] : cur(src, j);
        } else if (!ghost_row) {
            const int in = (i == 0) ? 1 : nx;
            const int side = (i == 0) ? 0 : 1;
            po[ix(i, j, pc)] = ghost(cur(in, j), a.bt[side][j - 1], a.bv[side][j - 1], g.dx,
                                     D(in, j));
        }
        return;
    }
    if (ghost_row) {
        const int in = (j == 0) ? 1 : ny;
        const int side = (j == 0) ? 2 : 3;
        po[ix(i, j, pc)] = ghost(cur(i, in), a.bt[side][i - 1], a.bv[side][i - 1], g.dy,
                                 D(i, in));
        return;
    }
    if (!a.has_solid || (fl[ix(i, j, pc)] & 1)) return;
    // interior solid cell: mean of the fluid neighbours (mirrored for a fixed value)
    double s = 0.0, c = 0.0;
    const int ni[4] = {i + 1, i - 1, i, i};
    const int nj[4] = {j, j, j + 1, j - 1};
    double pf[4], w[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const bool f = fl[ix(ni[q], nj[q], pc)] & 1;
        pf[q] = f ? po[ix(ni[q], nj[q], pc)] : 0.0;
        w[q] = f ? 1.0 : 0.0;
    }
    s = pf[0] + pf[1] + pf[2] + pf[3];
    c = w[0] + w[1] + w[2] + w[3];
    if (c > 0) {
        const double avg = s / fmax(c, 1.0);
        po[ix(i, j, pc)] = a.obs_t == 1 ? 2 * a.obs_v - avg : avg;
    } else if (a.obs_t == 1) {
        po[ix(i, j, pc)] = a.obs_v;
    }
}

// ---------------------------------------------------------------------------
// Per-step monitors and state advance (solver.py:378-401, 427-443, 517-540,
// models.py:210-250, 329-333, 421-426).  Partial sums per block go to
// `part`; the last block adds them in block order (deterministic), writes the
// record and advances t, n, dt_old.
// ---------------------------------------------------------------------------
static_assert(REC_NV + 4 < CFD_MAX_MON, "the record holds the sums, nu_t, force, VOF values");
// NS: the number of transported fields (a compile-time count keeps the
// per-thread sums and the block reductions as small as the case needs)
template <int NS>
__global__ void __launch_bounds__(RED_THREADS) k_record(RecordArgs a, Geo g) {
    constexpr int NV = 10 + 5 * NS;
    __shared__ double lds[NV * (RED_THREADS / CFD_WAVE)];
    const int nx = g.nx, ny = g.ny, pc = g.pc;
    double val[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) val[q] = 0.0;
    const bool active = !a.st->done;
    if (active) {
        const int tid = blockIdx.x * blockDim.x + threadIdx.x;
        const int nth = gridDim.x * blockDim.x;
        struct Cell {
            double uw, ue, vs, vn, ph[NS > 0 ? NS : 1];
            bool f;
        };
        flat_reduce<8>(
            nx * ny, ny, 1, 1,
            [&](int i, int j) {
                Cell e;
                e.f = a.fl[ix(i, j, pc)] & 1;
                e.uw = a.u[ix(i - 1, j, g.pu)];
                e.ue = a.u[ix(i, j, g.pu)];
                e.vs = a.v[ix(i, j - 1, g.pv)];
                e.vn = a.v[ix(i, j, g.pv)];
#pragma unroll
                for (int s = 0; s < NS; ++s) e.ph[s] = a.phi[s][ix(i, j, pc)];
                return e;
            },
            [&](const Cell& e, bool ok) {
                const bool f = ok && e.f;
                const double div = (e.ue - e.uw) / g.dx + (e.vn - e.vs) / g.dy;
                val[0] = f ? nanmax(val[0], fabs(div)) : val[0];
                const double uc = 0.5 * (e.uw + e.ue);
                const double vc = 0.5 * (e.vs + e.vn);
                val[1] = f ? val[1] + (uc * uc + vc * vc) : val[1];
                // compile-time scalar indices keep val[] in registers
#pragma unroll
                for (int s = 0; s < NS; ++s)
                    val[10 + 5 * s] = f ? val[10 + 5 * s] + e.ph[s] : val[10 + 5 * s];
            });
        if (a.has_solid) {
            // fluid cells next to interior solid cells (solver.py:542-567)
            for (int k = tid; k < nx * ny; k += nth) {
                const int i = 1 + k / ny, j = 1 + k % ny;
                if (!(a.fl[ix(i, j, pc)] & 1)) continue;
                auto solid = [&](int si, int sj) -> bool {
                    return si >= 1 && si <= nx && sj >= 1 && sj <= ny &&
                           !(a.fl[ix(si, sj, pc)] & 1);
                };
                const double uc = 0.5 * (a.u[ix(i - 1, j, g.pu)] + a.u[ix(i, j, g.pu)]);
                const double vc = 0.5 * (a.v[ix(i, j - 1, g.pv)] + a.v[ix(i, j, g.pv)]);
                const double pv = a.p[ix(i, j, pc)];
                if (solid(i + 1, j)) { val[2] += pv; val[8] += vc; }
                if (solid(i - 1, j)) { val[3] += pv; val[9] += vc; }
                if (solid(i, j + 1)) { val[4] += pv; val[6] += uc; }
                if (solid(i, j - 1)) { val[5] += pv; val[7] += uc; }
            }
        }
        // wall fluxes of scalars with fixed-value or fixed-flux sides (models.py:228-250)
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            if (!a.wallflux[s]) continue;
            const double* phi = a.phi[s];
#pragma unroll
            for (int side = 0; side < 4; ++side) {
                const int len = side < 2 ? ny : nx;
                const double h = side < 2 ? g.dx : g.dy;
                for (long k = tid; k < len; k += nth) {
                    int ci, cj;
                    if (side == 0) { ci = 1; cj = (int)k + 1; }
                    else if (side == 1) { ci = nx; cj = (int)k + 1; }
                    else if (side == 2) { ci = (int)k + 1; cj = 1; }
                    else { ci = (int)k + 1; cj = ny; }
                    const int code = a.bt[s][side][k];
                    const double bv = a.bv[s][side][k];
                    double q = 0.0;
                    if (code == 1) {
                        const double D = a.diffs[s] ? a.diffs[s][ix(ci, cj, pc)] : a.D0[s];
                        q = -D * (bv - phi[ix(ci, cj, pc)]) / (0.5 * h);
                    } else if (code == 2) {
                        q = -bv;
                    }
                    val[11 + 5 * s + side] += q;
                }
            }
        }
    }
    // block reduction: max for slot 0, sums elsewhere
    {
        double mx[1] = {val[0]};
        block_max<1>(mx, lds);
        val[0] = mx[0];
        double sm[NV - 1];
#pragma unroll
        for (int q = 1; q < NV; ++q) sm[q - 1] = val[q];
        block_sum<NV - 1>(sm, lds);
#pragma unroll
        for (int q = 1; q < NV; ++q) val[q] = sm[q - 1];
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) a.part[NV * blockIdx.x + q] = val[q];
    }
    if (!gather_block0(a.flags, a.spin)) return;
    // block 0: add the block partials in block order (fixed tree per thread)
    double tot[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) tot[q] = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) {
        tot[0] = nanmax(tot[0], a.part[NV * b]);
#pragma unroll
        for (int q = 1; q < NV; ++q) tot[q] += a.part[NV * b + q];
    }
    {
        double mx[1] = {tot[0]};
        block_max<1>(mx, lds);
        tot[0] = mx[0];
        double sm[NV - 1];
#pragma unroll
        for (int q = 1; q < NV; ++q) sm[q - 1] = tot[q];
        block_sum<NV - 1>(sm, lds);
#pragma unroll
        for (int q = 1; q < NV; ++q) tot[q] = sm[q - 1];
    }
    if (threadIdx.x != 0) return;
    DevState* st = a.st;
    if (!st->done) {
        const double dt = st->dt;
        st->has_old = a.scheme_ab2 ? 1 : 0;
        st->dt_old = dt;
        st->t += dt;
        st->n += 1;
        StepRecord r;
        r.t = st->t;
        r.dt = dt;
        r.gamma = st->gamma;
        r.residual = st->residual / dt;
        r.divergence = tot[0];
        r.ke = 0.5 * tot[1] * g.dx * g.dy;
        r.cg_relres = st->cg_relres;
        r.n = st->n;
        r.cg_iters = st->cg_iters;
        r.cg_info = st->cg_info;
        r.cg_outer = st->cg_outer;
        r.nmon = 10 + 5 * a.nscalar;
        for (int q = 0; q < CFD_MAX_MON; ++q) r.mon[q] = q < NV ? tot[q] : 0.0;
        r.mon[REC_NV] = st->nut_max;   // turbulence.py:158-159, 322-324
        r.mon[REC_NV + 1] = st->flow_force;   // solver.py:439-440
        r.mon[REC_NV + 2] = st->vof_sum;      // vof.py:264-281: alpha sum, front, height
        r.mon[REC_NV + 3] = st->vof_front;
        r.mon[REC_NV + 4] = st->vof_height;
        const double res = r.residual;
        if (!(res == res) || res > 1e8) {
            st->error = 1;          // solver.py:395-399: no record for this step
            st->done = 1;
            r.nmon = -1;
        }
        a.ring[(st->n - 1) % a.ring_len] = r;
        if (a.check_t_end && !(st->t < a.t_end - 1e-12)) st->done = 1;
        if (a.max_steps > 0 && st->n >= a.max_steps) st->done = 1;
        if (a.steady_tol > 0.0 && st->n > 10 && res < a.steady_tol) st->done = 1;
    }
}

// Reset the grid-barrier counter before a persistent launch (a kernel, not a
// memset, so the sequence stays valid inside captured graphs).
__global__ void k_zero_u32(unsigned* p, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0u;
}
