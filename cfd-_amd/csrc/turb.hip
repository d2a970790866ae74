// Turbulence models on the device (cfd2d/turbulence.py): the Smagorinsky LES
// model and Menter's k-omega SST (2003, low-Re form).  Per step:
//   k_turb_cells   before the time-step control: the strain rate S^2 and the
//                  turbulent viscosity nu_t of every inner cell; SST also the
//                  blending function F1, the cross-diffusion term and the
//                  blended diffusivities of k and omega (pre_momentum,
//                  turbulence.py:151-156, 250-298)
//   k_turb_ghosts  ghost cells of nu_t (turbulence.py:109-124) and of the
//                  two diffusivities (turbulence.py:294-298)
//   k_sst_update   after the pressure correction: transport of k and omega
//                  (the TVD face fluxes of the passive scalars with the
//                  variable diffusivities) with their sources treated
//                  semi-implicitly and the lower clips (models.py:186-201,
//                  turbulence.py:301-320)
//   k_sst_wall     the viscous omega of the cells next to a wall
//                  (turbulence.py:242-246), between two ghost fills as in
//                  the reference.
// Every expression keeps the reference's operand order; numpy's maximum and
// minimum (NaN-propagating) are mirrored by np_max/np_min.  tanh and pow of
// the SST blending come from the device math library, so SST agrees with the
// reference to rounding, not bit for bit; the Smagorinsky path uses only
// sqrt and is bit-exact.
#pragma clang fp contract(off)
#include "cfd_kernels.h"

enum : int { TURB_NONE = 0, TURB_SMAGORINSKY = 1, TURB_SST = 2 };

namespace sst {
constexpr double a1 = 0.31, beta_star = 0.09;
constexpr double sk1 = 0.85, so1 = 0.5, beta1 = 0.075, gamma1 = 5.0 / 9.0;
constexpr double sk2 = 1.0, so2 = 0.856, beta2 = 0.0828, gamma2 = 0.44;
}  // namespace sst

__device__ __forceinline__ double np_max(double a, double b) { return (a >= b || a != a) ? a : b; }
__device__ __forceinline__ double np_min(double a, double b) { return (a <= b || a != a) ? a : b; }

struct TurbArgs {
    int model;
    double nu;
    const double* u;
    const double* v;
    const uint8_t* fl;
    double* nut;             // (nx+2) x P
    const double* l2;        // Smagorinsky: min(C_s Delta, kappa d)^2 per cell
    const double* kphi;      // SST: current k and omega (ghosts filled)
    const double* wphi;
    const double* d;         // wall distance, at least 1e-12
    double* F1;
    double* cross;           // grad k . grad omega
    double* dk;              // diffusivities of k and omega (with ghosts)
    double* dw;
    int kind[4];             // BC kind per side (left, right, bottom, top)
    const DevState* st;      // nullptr at set-up
};

// S^2 = 2 S_ij S_ij at the inner cell (i, j) (turbulence.py:88-99): the
// normal strains of the cell and the shear of its four nodes
__device__ __forceinline__ double strain2(const double* __restrict__ u,
                                          const double* __restrict__ v, const Geo& g, int i,
                                          int j) {
    const double dudx = (u[ix(i, j, g.pu)] - u[ix(i - 1, j, g.pu)]) / g.dx;
    const double dvdy = (v[ix(i, j, g.pv)] - v[ix(i, j - 1, g.pv)]) / g.dy;
    auto shear = [&](int a, int b) {   // node (a, b): du/dy + dv/dx
        return (u[ix(a, b + 1, g.pu)] - u[ix(a, b, g.pu)]) / g.dy +
               (v[ix(a + 1, b, g.pv)] - v[ix(a, b, g.pv)]) / g.dx;
    };
    const double s00 = shear(i - 1, j - 1), s10 = shear(i, j - 1);
    const double s01 = shear(i - 1, j), s11 = shear(i, j);
    const double sh2 = 0.25 * (s00 * s00 + s10 * s10 + s01 * s01 + s11 * s11);
    return 2.0 * (dudx * dudx + dvdy * dvdy) + sh2;
}

// one thread per inner cell
__global__ void __launch_bounds__(256) k_turb_cells(TurbArgs a, Geo g) {
    const int2 rc_ = xcd_rowgrid();
    const int j = rc_.x * blockDim.x + threadIdx.x + 1;
    const int i = rc_.y + 1;
    if (j > g.ny) return;
    if (a.st && a.st->done) return;
    const int c = ix(i, j, g.pc);
    const bool fluid = a.fl[c] & 1;
    const double S2 = strain2(a.u, a.v, g, i, j);
    if (a.model == TURB_SMAGORINSKY) {
        a.nut[c] = fluid ? a.l2[c] * sqrt(S2) : 0.0;
        return;
    }
    using namespace sst;
    const double nu = a.nu;
    const double k = np_max(a.kphi[c], 0.0), w = np_max(a.wphi[c], 1e-12);
    const int P = g.pc;
    const double dkx = (a.kphi[c + P] - a.kphi[c - P]) / (2 * g.dx);
    const double dky = (a.kphi[c + 1] - a.kphi[c - 1]) / (2 * g.dy);
    const double dwx = (a.wphi[c + P] - a.wphi[c - P]) / (2 * g.dx);
    const double dwy = (a.wphi[c + 1] - a.wphi[c - 1]) / (2 * g.dy);
    const double cross = dkx * dwx + dky * dwy;
    const double CD = np_max(2.0 * so2 / w * cross, 1e-10);
    const double sk = sqrt(k);
    const double d = a.d[c];
    const double d2 = d * d;
    const double visc = 500 * nu / (d2 * w);
    const double arg1 = np_min(np_max(sk / (beta_star * w * d), visc), 4 * so2 * k / (CD * d2));
    const double F1 = tanh(pow(arg1, 4.0));
    const double arg2 = np_max(2 * sk / (beta_star * w * d), visc);
    const double F2 = tanh(arg2 * arg2);
    const double nut = a1 * k / np_max(a1 * w, sqrt(S2) * F2);
    a.nut[c] = fluid ? nut : 0.0;
    a.F1[c] = F1;
    a.cross[c] = cross;
    a.dk[c] = nu + (F1 * sk1 + (1 - F1) * sk2) * nut;
    a.dw[c] = nu + (F1 * so1 + (1 - F1) * so2) * nut;
}

// ghost cells: the reference fills the ghost columns over the full height,
// then the ghost rows over the full width from the row next to them, so a
// corner takes the value of its diagonal inner neighbour's column copy.
// nt: nu_t, zero next to no-slip walls (turbulence.py:109-124); the
// diffusivities copy their neighbour (turbulence.py:294-298).
__global__ void __launch_bounds__(256) k_turb_ghosts(TurbArgs a, Geo g) {
    const int nx = g.nx, ny = g.ny, P = g.pc;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    // t < 2 (nx + 2): rows j = 0 / ny+1; then columns i = 0 / nx+1 for j = 1..ny
    int i, j;
    if (t < 2 * (nx + 2)) {
        i = t % (nx + 2);
        j = t < nx + 2 ? 0 : ny + 1;
    } else if (t < 2 * (nx + 2) + 2 * ny) {
        const int r = t - 2 * (nx + 2);
        j = 1 + r % ny;
        i = r < ny ? 0 : nx + 1;
    } else {
        return;
    }
    if (a.st && a.st->done) return;
    // column pass value of a ghost column cell (rows 1..ny)
    auto col_nut = [&](int ci, int cj) -> double {
        if (g.periodic) return a.nut[ix(ci == 0 ? nx : 1, cj, P)];
        const int side = ci == 0 ? 0 : 1;
        return a.kind[side] == BC_NOSLIP ? 0.0 : a.nut[ix(ci == 0 ? 1 : nx, cj, P)];
    };
    auto col_d = [&](const double* f, int ci, int cj) -> double {
        return f[ix(ci == 0 ? 1 : (ci == nx + 1 ? nx : ci), cj, P)];
    };
    double nt, vk = 0.0, vw = 0.0;
    if (j == 0 || j == ny + 1) {
        const int side = j == 0 ? 2 : 3;
        const int jn = j == 0 ? 1 : ny;
        const bool inner = i >= 1 && i <= nx;
        nt = a.kind[side] == BC_NOSLIP ? 0.0 : (inner ? a.nut[ix(i, jn, P)] : col_nut(i, jn));
        if (a.model == TURB_SST) {
            vk = col_d(a.dk, i, jn);
            vw = col_d(a.dw, i, jn);
        }
    } else {
        nt = col_nut(i, j);
        if (a.model == TURB_SST) {
            vk = col_d(a.dk, i, j);
            vw = col_d(a.dw, i, j);
        }
    }
    const int c = ix(i, j, P);
    a.nut[c] = nt;
    if (a.model == TURB_SST) {
        a.dk[c] = vk;
        a.dw[c] = vw;
    }
}

struct SstArgs {
    ScalarArgs k, w;         // transport of k and omega (phi_in -> phi_out, diff set)
    const double* nut;
    const double* F1;
    const double* cross;
};

// k and omega after the step (models.py:186-201 with the SST sources of
// turbulence.py:301-320): explicit transport, then
// phi = (phi + dt S_exp) / (1 + dt max(S_imp, 0)), clipped below; the strain
// rate is that of the corrected velocity (turbulence.py:319).  Both sources
// use the k and omega of the start of the step.
__global__ void __launch_bounds__(256) k_sst_update(SstArgs a, Geo g) {
    using namespace sst;
    const int2 rc_ = xcd_rowgrid();
    const int j = rc_.x * blockDim.x + threadIdx.x;
    const int i = rc_.y;
    const int nx = g.nx, ny = g.ny, pc = g.pc;
    if (j > ny + 1) return;
    if (a.k.st->done) return;
    const int c = ix(i, j, pc);
    const bool inner = i >= 1 && i <= nx && j >= 1 && j <= ny;
    const bool fluid = inner && (a.k.fl[c] & 1);
    const double k0 = a.k.phi_in[c], w0 = a.w.phi_in[c];
    if (!fluid) {
        a.k.phi_out[c] = k0;
        a.w.phi_out[c] = w0;
        return;
    }
    const double dt = a.k.st->dt;
    const int fl_face = (g.periodic && i == 1) ? nx : i - 1;
    double rk = xface_term<false>(a.k, g, fl_face, j, i);
    rk += xface_term<false>(a.k, g, i, j, i);
    rk += yface_term<false>(a.k, g, i, j - 1, j);
    rk += yface_term<false>(a.k, g, i, j, j);
    double rw = xface_term<false>(a.w, g, fl_face, j, i);
    rw += xface_term<false>(a.w, g, i, j, i);
    rw += yface_term<false>(a.w, g, i, j - 1, j);
    rw += yface_term<false>(a.w, g, i, j, j);
    const double S2 = strain2(a.k.u, a.k.v, g, i, j);
    const double kk = np_max(k0, 0.0), ww = np_max(w0, 1e-12);
    const double nut = a.nut[c];
    // k: Pk = min(nu_t S^2, 10 beta* k omega), sink beta* omega
    const double Pk = np_min(nut * S2, 10.0 * beta_star * kk * ww);
    double kn = k0 + dt * rk;
    kn = (kn + dt * Pk) / (1.0 + dt * np_max(beta_star * ww, 0.0));
    kn = np_max(kn, 1e-14);
    // omega: gamma Pk / nu_t (gamma S^2 where nu_t vanishes), cross diffusion
    const double F1 = a.F1[c], cr = a.cross[c];
    const double gam = F1 * gamma1 + (1 - F1) * gamma2;
    const double beta = F1 * beta1 + (1 - F1) * beta2;
    const double Pw = nut > 1e-14 ? gam * Pk / np_max(nut, 1e-14) : gam * S2;
    const double CDw = 2.0 * (1 - F1) * so2 / ww * cr;
    const double s_exp = Pw + np_max(CDw, 0.0);
    const double s_imp = beta * ww + np_max(-CDw, 0.0) / ww;
    double wn = w0 + dt * rw;
    wn = (wn + dt * s_exp) / (1.0 + dt * np_max(s_imp, 0.0));
    wn = np_max(wn, 1e-8);
    a.k.phi_out[c] = kn;
    a.w.phi_out[c] = wn;
}

// Between the two ghost fills of omega: keep the state of the first fill in
// `keep` (the second fill reads solid cells from it, as the reference's
// second fill_ghosts sees the first one's solid values) and set the wall
// cells' viscous omega in `w`.
__global__ void __launch_bounds__(256) k_sst_wall(double* __restrict__ w,
                                                  double* __restrict__ keep,
                                                  const uint8_t* __restrict__ wadj,
                                                  const double* __restrict__ wwall, Geo g,
                                                  const DevState* st) {
    const int2 rc_ = xcd_rowgrid();
    const int j = rc_.x * blockDim.x + threadIdx.x;
    const int i = rc_.y;
    if (j > g.ny + 1) return;
    if (st && st->done) return;
    const int c = ix(i, j, g.pc);
    const double x = w[c];
    keep[c] = x;
    if (wadj[c]) w[c] = wwall[c];
}
