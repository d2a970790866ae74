/*
 * CPU FP64 oracle: serial C restatement of the reference's Numba kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links or calls this file;
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * the shared library built from it (oracle/Makefile), as the checker.
 *
 * Array layout is the reference's: C order, index [i, j] with j contiguous,
 *   u, F, hu : (nx+1) x (ny+2)      v, G, hv : (nx+2) x (ny+1)
 *   p, phi, fluid, diff : (nx+2) x (ny+2)
 * Built with -ffp-contract=off so every product and sum rounds exactly as the
 * reference's IEEE double arithmetic does (Python floats / Numba without
 * fastmath).
 *
 * Follows cfd2d/kernels.py:
 *   oracle_momentum_tendencies  kernels.py:44-155 (constant-viscosity branch,
 *                               lines 104-105; no variable density)
 *   oracle_momentum_tendencies_var  the same with the stress-form diffusion of
 *                               a cell viscosity field (kernels.py:96-103,
 *                               131-138), divided by the face density when
 *                               rho is given (variable density, lines 104-105
 *                               and 139-140)
 *   oracle_vof_advect           vof.py:35-133 (van Leer + interface compression)
 *   oracle_obstacle_bc          kernels.py:158-209
 *   limiter / face_value        kernels.py:215-251
 *   oracle_scalar_transport     kernels.py:254-391
 *   oracle_pcg_csr              scipy.sparse.linalg.cg (SciPy 1.15
 *                               iterative.py:305-423) with the Jacobi
 *                               preconditioner of pressure.py:138-139, for the
 *                               all-cores CPU baseline
 *
 * Built with OpenMP: the momentum loops are gather-form (every face is written
 * once from read-only inputs), so their results do not depend on the thread
 * count; the scatter-form scalar transport stays serial.  The CG's dot
 * products are OpenMP reductions, summed in a thread-count-dependent order
 * (its results agree with SciPy's to rounding, not bit for bit).
 */
#include <math.h>
#include <stdint.h>

#define U_(i, j) u[(long)(i) * (ny + 2) + (j)]
#define V_(i, j) v[(long)(i) * (ny + 1) + (j)]
#define C_(a, i, j) a[(long)(i) * (ny + 2) + (j)]

void oracle_momentum_tendencies(const double *u, const double *v, const uint8_t *fluid,
                                int nx, int ny, double dx, double dy, double nu0, double gamma,
                                int periodic, double *hu, double *hv)
{
    const double idx2 = 1.0 / (dx * dx), idy2 = 1.0 / (dy * dy);
    const double idx = 1.0 / dx, idy = 1.0 / dy;
    for (long k = 0; k < (long)(nx + 1) * (ny + 2); ++k) hu[k] = 0.0;
    for (long k = 0; k < (long)(nx + 2) * (ny + 1); ++k) hv[k] = 0.0;
    const int imax = periodic ? nx : nx - 1;

    /* u faces: between cells i and i+1 */
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= imax; ++i) {
        const int ip = i < nx ? i + 1 : 1;
        for (int j = 1; j <= ny; ++j) {
            if (!(C_(fluid, i, j) && C_(fluid, i + 1, j))) continue;
            const double uc = U_(i, j), ue = U_(ip, j), uw = U_(i - 1, j);
            const double un = U_(i, j + 1), us = U_(i, j - 1);
            const double diff = nu0 * ((ue - 2.0 * uc + uw) * idx2 + (un - 2.0 * uc + us) * idy2);
            const double a_e = 0.5 * (uc + ue), a_w = 0.5 * (uw + uc);
            const double du2dx = idx * (a_e * a_e - a_w * a_w)
                + gamma * idx * (fabs(a_e) * 0.5 * (uc - ue) - fabs(a_w) * 0.5 * (uw - uc));
            const double v_n = 0.5 * (V_(i, j) + V_(i + 1, j));
            const double v_s = 0.5 * (V_(i, j - 1) + V_(i + 1, j - 1));
            const double duvdy = idy * (v_n * 0.5 * (uc + un) - v_s * 0.5 * (us + uc))
                + gamma * idy * (fabs(v_n) * 0.5 * (uc - un) - fabs(v_s) * 0.5 * (us - uc));
            hu[(long)i * (ny + 2) + j] = diff - du2dx - duvdy;
        }
    }
    /* v faces: between cells j and j+1 */
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= nx; ++i) {
        for (int j = 1; j < ny; ++j) {
            if (!(C_(fluid, i, j) && C_(fluid, i, j + 1))) continue;
            const double vc = V_(i, j), ve = V_(i + 1, j), vw = V_(i - 1, j);
            const double vn = V_(i, j + 1), vs = V_(i, j - 1);
            const double diff = nu0 * ((ve - 2.0 * vc + vw) * idx2 + (vn - 2.0 * vc + vs) * idy2);
            const double b_n = 0.5 * (vc + vn), b_s = 0.5 * (vs + vc);
            const double dv2dy = idy * (b_n * b_n - b_s * b_s)
                + gamma * idy * (fabs(b_n) * 0.5 * (vc - vn) - fabs(b_s) * 0.5 * (vs - vc));
            const double u_e = 0.5 * (U_(i, j) + U_(i, j + 1));
            const double u_w = 0.5 * (U_(i - 1, j) + U_(i - 1, j + 1));
            const double duvdx = idx * (u_e * 0.5 * (vc + ve) - u_w * 0.5 * (vw + vc))
                + gamma * idx * (fabs(u_e) * 0.5 * (vc - ve) - fabs(u_w) * 0.5 * (vw - vc));
            hv[(long)i * (ny + 1) + j] = diff - duvdx - dv2dy;
        }
    }
}

/* Stress-form diffusion div(nu (grad u + grad u^T)) with the cell field nu
   (nu_eff), kernels.py:96-103 for u faces and 131-138 for v faces; the
   convection terms are those of the constant branch. */
void oracle_momentum_tendencies_var(const double *u, const double *v, const uint8_t *fluid,
                                    int nx, int ny, double dx, double dy, const double *nu,
                                    double gamma, int periodic, double *hu, double *hv,
                                    const double *rho)
{
    const double idx2 = 1.0 / (dx * dx), idy2 = 1.0 / (dy * dy);
    const double idx = 1.0 / dx, idy = 1.0 / dy;
    for (long k = 0; k < (long)(nx + 1) * (ny + 2); ++k) hu[k] = 0.0;
    for (long k = 0; k < (long)(nx + 2) * (ny + 1); ++k) hv[k] = 0.0;
    const int imax = periodic ? nx : nx - 1;
    for (int i = 1; i <= imax; ++i) {
        const int ip = i < nx ? i + 1 : 1;
        for (int j = 1; j <= ny; ++j) {
            if (!(C_(fluid, i, j) && C_(fluid, i + 1, j))) continue;
            const double uc = U_(i, j), ue = U_(ip, j), uw = U_(i - 1, j);
            const double un = U_(i, j + 1), us = U_(i, j - 1);
            const double nu_e = C_(nu, i + 1, j), nu_w = C_(nu, i, j);
            const double nu_n = 0.25 * (C_(nu, i, j) + C_(nu, i + 1, j) + C_(nu, i, j + 1)
                                        + C_(nu, i + 1, j + 1));
            const double nu_s = 0.25 * (C_(nu, i, j) + C_(nu, i + 1, j) + C_(nu, i, j - 1)
                                        + C_(nu, i + 1, j - 1));
            double diff = (2.0 * nu_e * (ue - uc) - 2.0 * nu_w * (uc - uw)) * idx2
                + idy * (nu_n * ((un - uc) * idy + (V_(i + 1, j) - V_(i, j)) * idx)
                         - nu_s * ((uc - us) * idy + (V_(i + 1, j - 1) - V_(i, j - 1)) * idx));
            if (rho) diff /= 0.5 * (C_(rho, i, j) + C_(rho, i + 1, j));
            const double a_e = 0.5 * (uc + ue), a_w = 0.5 * (uw + uc);
            const double du2dx = idx * (a_e * a_e - a_w * a_w)
                + gamma * idx * (fabs(a_e) * 0.5 * (uc - ue) - fabs(a_w) * 0.5 * (uw - uc));
            const double v_n = 0.5 * (V_(i, j) + V_(i + 1, j));
            const double v_s = 0.5 * (V_(i, j - 1) + V_(i + 1, j - 1));
            const double duvdy = idy * (v_n * 0.5 * (uc + un) - v_s * 0.5 * (us + uc))
                + gamma * idy * (fabs(v_n) * 0.5 * (uc - un) - fabs(v_s) * 0.5 * (us - uc));
            hu[(long)i * (ny + 2) + j] = diff - du2dx - duvdy;
        }
    }
    for (int i = 1; i <= nx; ++i) {
        for (int j = 1; j < ny; ++j) {
            if (!(C_(fluid, i, j) && C_(fluid, i, j + 1))) continue;
            const double vc = V_(i, j), ve = V_(i + 1, j), vw = V_(i - 1, j);
            const double vn = V_(i, j + 1), vs = V_(i, j - 1);
            const double nu_n = C_(nu, i, j + 1), nu_s = C_(nu, i, j);
            const double nu_e = 0.25 * (C_(nu, i, j) + C_(nu, i + 1, j) + C_(nu, i, j + 1)
                                        + C_(nu, i + 1, j + 1));
            const double nu_w = 0.25 * (C_(nu, i, j) + C_(nu, i - 1, j) + C_(nu, i, j + 1)
                                        + C_(nu, i - 1, j + 1));
            double diff = (2.0 * nu_n * (vn - vc) - 2.0 * nu_s * (vc - vs)) * idy2
                + idx * (nu_e * ((ve - vc) * idx + (U_(i, j + 1) - U_(i, j)) * idy)
                         - nu_w * ((vc - vw) * idx + (U_(i - 1, j + 1) - U_(i - 1, j)) * idy));
            if (rho) diff /= 0.5 * (C_(rho, i, j) + C_(rho, i, j + 1));
            const double b_n = 0.5 * (vc + vn), b_s = 0.5 * (vs + vc);
            const double dv2dy = idy * (b_n * b_n - b_s * b_s)
                + gamma * idy * (fabs(b_n) * 0.5 * (vc - vn) - fabs(b_s) * 0.5 * (vs - vc));
            const double u_e = 0.5 * (U_(i, j) + U_(i, j + 1));
            const double u_w = 0.5 * (U_(i - 1, j) + U_(i - 1, j + 1));
            const double duvdx = idx * (u_e * 0.5 * (vc + ve) - u_w * 0.5 * (vw + vc))
                + gamma * idx * (fabs(u_e) * 0.5 * (vc - ve) - fabs(u_w) * 0.5 * (vw - vc));
            hv[(long)i * (ny + 1) + j] = diff - duvdx - dv2dy;
        }
    }
}

void oracle_obstacle_bc(double *u, double *v, const uint8_t *fluid, int nx, int ny, int periodic)
{
    const int imax = periodic ? nx : nx - 1;
    for (int i = 1; i <= imax; ++i) {
        for (int j = 1; j <= ny; ++j) {
            const int fl = C_(fluid, i, j), fr = C_(fluid, i + 1, j);
            if (fl && fr) continue;
            if (fl || fr) { U_(i, j) = 0.0; continue; }
            double s = 0.0;
            int c = 0;
            if (C_(fluid, i, j + 1) && C_(fluid, i + 1, j + 1)) { s += U_(i, j + 1); ++c; }
            if (C_(fluid, i, j - 1) && C_(fluid, i + 1, j - 1)) { s += U_(i, j - 1); ++c; }
            U_(i, j) = c > 0 ? -s / c : 0.0;
        }
    }
    if (periodic)
        for (int j = 0; j < ny + 2; ++j) U_(0, j) = U_(nx, j);
    for (int i = 1; i <= nx; ++i) {
        for (int j = 1; j < ny; ++j) {
            const int fb = C_(fluid, i, j), ft = C_(fluid, i, j + 1);
            if (fb && ft) continue;
            if (fb || ft) { V_(i, j) = 0.0; continue; }
            double s = 0.0;
            int c = 0;
            if (C_(fluid, i + 1, j) && C_(fluid, i + 1, j + 1)) { s += V_(i + 1, j); ++c; }
            if (C_(fluid, i - 1, j) && C_(fluid, i - 1, j + 1)) { s += V_(i - 1, j); ++c; }
            V_(i, j) = c > 0 ? -s / c : 0.0;
        }
    }
}

double oracle_limiter(double r, int kind)
{
    switch (kind) {
    case 0: return 0.0;
    case 1: return r <= 0.0 ? 0.0 : 2.0 * r / (1.0 + r);
    case 2: return r <= 0.0 ? 0.0 : fmin(r, 1.0);
    case 3: return r <= 0.0 ? 0.0 : fmax(fmin(2.0 * r, 1.0), fmin(r, 2.0));
    case 5: return r <= 0.0 ? 0.0 : fmin(2.0 * r, fmin(0.5 * (1.0 + r), 2.0));
    default: return 1.0;
    }
}

static double face_value(double phi_u, double phi_c, double phi_d, int has_u, int kind)
{
    if (kind == 0) return phi_c;
    if (kind == 4) return 0.5 * (phi_c + phi_d);
    if (!has_u) return phi_c;
    const double den = phi_d - phi_c;
    if (fabs(den) < 1e-300) return phi_c;
    const double r = (phi_c - phi_u) / den;
    return phi_c + 0.5 * oracle_limiter(r, kind) * den;
}

void oracle_scalar_transport(const double *phi, const double *u, const double *v,
                             const uint8_t *fluid, int nx, int ny, double dx, double dy,
                             const double *diff, int limiter, int periodic,
                             const int64_t *bt_l, const double *bv_l,
                             const int64_t *bt_r, const double *bv_r,
                             const int64_t *bt_b, const double *bv_b,
                             const int64_t *bt_t, const double *bv_t,
                             int obs_t, double obs_v, double *out)
{
    for (long k = 0; k < (long)(nx + 2) * (ny + 2); ++k) out[k] = 0.0;

    for (int i = 0; i <= nx; ++i) {
        for (int j = 1; j <= ny; ++j) {
            if (i == 0 || (i == nx && !periodic)) {
                int P, t;
                double sgn, val;
                if (i == 0) {
                    if (periodic) continue;
                    P = 1; sgn = 1.0; t = (int)bt_l[j - 1]; val = bv_l[j - 1];
                } else {
                    P = nx; sgn = -1.0; t = (int)bt_r[j - 1]; val = bv_r[j - 1];
                }
                if (!C_(fluid, P, j)) continue;
                const double un = sgn * U_(i, j);
                const double phb = (t == 1 && un > 0.0) ? val : C_(phi, P, j);
                double flux_in = un * phb;
                if (t == 1) flux_in += C_(diff, P, j) * (val - C_(phi, P, j)) / (0.5 * dx);
                else if (t == 2) flux_in += val;
                C_(out, P, j) += flux_in / dx;
                continue;
            }
            const int L = i, R = i < nx ? i + 1 : 1;
            const int fL = C_(fluid, L, j), fR = C_(fluid, R, j);
            if (fL && fR) {
                const double uf = U_(i, j);
                int Cc, D, Ui;
                if (uf >= 0.0) {
                    Cc = L; D = R; Ui = L - 1;
                    if (Ui < 1) Ui = periodic ? nx : -1;
                } else {
                    Cc = R; D = L; Ui = R + 1;
                    if (Ui > nx) Ui = periodic ? 1 : -1;
                }
                const int has_u = Ui >= 1 && C_(fluid, Ui, j);
                const double phu = has_u ? C_(phi, Ui, j) : 0.0;
                const double phf = face_value(phu, C_(phi, Cc, j), C_(phi, D, j), has_u, limiter);
                const double dl = C_(diff, L, j), dr = C_(diff, R, j);
                const double df = dl + dr > 0.0 ? 2.0 * dl * dr / (dl + dr) : 0.0;
                const double F = uf * phf - df * (C_(phi, R, j) - C_(phi, L, j)) / dx;
                C_(out, L, j) -= F / dx;
                C_(out, R, j) += F / dx;
            } else if ((fL || fR) && obs_t == 1) {
                const int P = fL ? L : R;
                C_(out, P, j) += C_(diff, P, j) * (obs_v - C_(phi, P, j)) / (0.5 * dx) / dx;
            }
        }
    }

    for (int i = 1; i <= nx; ++i) {
        for (int j = 0; j <= ny; ++j) {
            if (j == 0 || j == ny) {
                int P, t;
                double sgn, val;
                if (j == 0) { P = 1; sgn = 1.0; t = (int)bt_b[i - 1]; val = bv_b[i - 1]; }
                else { P = ny; sgn = -1.0; t = (int)bt_t[i - 1]; val = bv_t[i - 1]; }
                if (!C_(fluid, i, P)) continue;
                const double vn = sgn * V_(i, j);
                const double phb = (t == 1 && vn > 0.0) ? val : C_(phi, i, P);
                double flux_in = vn * phb;
                if (t == 1) flux_in += C_(diff, i, P) * (val - C_(phi, i, P)) / (0.5 * dy);
                else if (t == 2) flux_in += val;
                C_(out, i, P) += flux_in / dy;
                continue;
            }
            const int B = j, T = j + 1;
            const int fB = C_(fluid, i, B), fT = C_(fluid, i, T);
            if (fB && fT) {
                const double vf = V_(i, j);
                int Cc, D, Uj;
                if (vf >= 0.0) { Cc = B; D = T; Uj = B - 1; }
                else { Cc = T; D = B; Uj = T + 1; }
                const int has_u = 1 <= Uj && Uj <= ny && C_(fluid, i, Uj);
                const double phu = has_u ? C_(phi, i, Uj) : 0.0;
                const double phf = face_value(phu, C_(phi, i, Cc), C_(phi, i, D), has_u, limiter);
                const double db = C_(diff, i, B), dt_ = C_(diff, i, T);
                const double df = db + dt_ > 0.0 ? 2.0 * db * dt_ / (db + dt_) : 0.0;
                const double F = vf * phf - df * (C_(phi, i, T) - C_(phi, i, B)) / dy;
                C_(out, i, B) -= F / dy;
                C_(out, i, T) += F / dy;
            } else if ((fB || fT) && obs_t == 1) {
                const int P = fB ? B : T;
                C_(out, i, P) += C_(diff, i, P) * (obs_v - C_(phi, i, P)) / (0.5 * dy) / dy;
            }
        }
    }

    for (int i = 0; i < nx + 2; ++i)
        for (int j = 0; j < ny + 2; ++j)
            if (!C_(fluid, i, j) || i == 0 || i == nx + 1 || j == 0 || j == ny + 1)
                C_(out, i, j) = 0.0;
}

/* vof.py:22-31: van Leer face value from the upwind, central and downwind cells */
static double vl_face(double pu, double pc, double pd, int has_u)
{
    if (!has_u) return pc;
    const double den = pd - pc;
    if (fabs(den) < 1e-300) return pc;
    const double r = (pc - pu) / den;
    const double psi = r <= 0.0 ? 0.0 : 2.0 * r / (1.0 + r);
    return pc + 0.5 * psi * den;
}

/* vof.py:34-133: one explicit step of alpha; faces in the reference's loop
   order, each adding to and taking from its two cells; inflow faces bring
   a_in; the compression flux u_r a_up (1 - a_down) with u_r = c_alpha |u| n;
   alpha clipped to [0, 1] on fluid cells at the end. */
void oracle_vof_advect(const double *a, const double *u, const double *v, const uint8_t *fluid,
                       int nx, int ny, double dx, double dy, double dt, double c_alpha,
                       int periodic, double a_in, double *out)
{
    for (long k = 0; k < (long)(nx + 2) * (ny + 2); ++k) out[k] = a[k];
    for (int i = 0; i <= nx; ++i)
        for (int j = 1; j <= ny; ++j) {
            const double uf = U_(i, j);
            if (i == 0 || (i == nx && !periodic)) {
                if (i == 0 && periodic) continue;
                const int P = i == 0 ? 1 : nx;
                if (!C_(fluid, P, j)) continue;
                const double into = i == 0 ? uf : -uf;
                const double ab = into > 0 ? a_in : C_(a, P, j);
                C_(out, P, j) += dt * into * ab / dx;
                continue;
            }
            const int L = i, R = i < nx ? i + 1 : 1;
            if (!(C_(fluid, L, j) && C_(fluid, R, j))) continue;
            int Cc, D, Ui;
            if (uf >= 0.0) {
                Cc = L; D = R; Ui = L - 1;
                if (Ui < 1) Ui = periodic ? nx : -1;
            } else {
                Cc = R; D = L; Ui = R + 1;
                if (Ui > nx) Ui = periodic ? 1 : -1;
            }
            const int has_u = Ui >= 1 && C_(fluid, Ui, j);
            const double af = vl_face(has_u ? C_(a, Ui, j) : 0.0, C_(a, Cc, j), C_(a, D, j), has_u);
            double flux = uf * af;
            const double gx = (C_(a, R, j) - C_(a, L, j)) / dx;
            const int jn = j < ny ? j + 1 : j, js = j > 1 ? j - 1 : j;
            const double gy = 0.25 * ((C_(a, L, jn) - C_(a, L, js)) + (C_(a, R, jn) - C_(a, R, js))) / dy;
            const double gm = sqrt(gx * gx + gy * gy);
            if (gm > 1e-12 && c_alpha > 0.0) {
                const double ur = c_alpha * fabs(uf) * gx / gm;
                const double ac = ur >= 0.0 ? C_(a, L, j) : C_(a, R, j);
                const double ad = ur >= 0.0 ? C_(a, R, j) : C_(a, L, j);
                flux += ur * ac * (1.0 - ad);
            }
            C_(out, L, j) -= dt * flux / dx;
            C_(out, R, j) += dt * flux / dx;
        }
    for (int i = 1; i <= nx; ++i)
        for (int j = 0; j <= ny; ++j) {
            const double vf = V_(i, j);
            if (j == 0 || j == ny) {
                const int P = j == 0 ? 1 : ny;
                if (!C_(fluid, i, P)) continue;
                const double into = j == 0 ? vf : -vf;
                const double ab = into > 0 ? a_in : C_(a, i, P);
                C_(out, i, P) += dt * into * ab / dy;
                continue;
            }
            const int B = j, T = j + 1;
            if (!(C_(fluid, i, B) && C_(fluid, i, T))) continue;
            int Cc, D, Uj;
            if (vf >= 0.0) { Cc = B; D = T; Uj = B - 1; }
            else { Cc = T; D = B; Uj = T + 1; }
            const int has_u = Uj >= 1 && Uj <= ny && C_(fluid, i, Uj);
            const double af = vl_face(has_u ? C_(a, i, Uj) : 0.0, C_(a, i, Cc), C_(a, i, D), has_u);
            double flux = vf * af;
            const double gy = (C_(a, i, T) - C_(a, i, B)) / dy;
            int ie = (i < nx || periodic) ? i + 1 : i;
            int iw = (i > 1 || periodic) ? i - 1 : i;
            if (ie > nx) ie = 1;
            if (iw < 1) iw = nx;
            const double gx = 0.25 * ((C_(a, ie, B) - C_(a, iw, B)) + (C_(a, ie, T) - C_(a, iw, T))) / dx;
            const double gm = sqrt(gx * gx + gy * gy);
            if (gm > 1e-12 && c_alpha > 0.0) {
                const double ur = c_alpha * fabs(vf) * gy / gm;
                const double ac = ur >= 0.0 ? C_(a, i, B) : C_(a, i, T);
                const double ad = ur >= 0.0 ? C_(a, i, T) : C_(a, i, B);
                flux += ur * ac * (1.0 - ad);
            }
            C_(out, i, B) -= dt * flux / dy;
            C_(out, i, T) += dt * flux / dy;
        }
    for (int i = 1; i <= nx; ++i)
        for (int j = 1; j <= ny; ++j)
            if (C_(fluid, i, j)) {
                if (C_(out, i, j) < 0.0) C_(out, i, j) = 0.0;
                else if (C_(out, i, j) > 1.0) C_(out, i, j) = 1.0;
            }
}

/* Jacobi-preconditioned CG on a CSR matrix: the loop of
   scipy.sparse.linalg.cg (SciPy 1.15 iterative.py:305-423) -- stop when
   ||r|| < rtol ||b|| at the top of an iteration, z = r / diag, p = z (first
   iteration) or z + beta p, alpha = rho / (p.q) -- with OpenMP loops and
   reductions.  x holds x0 on entry and the result on return; *iters the
   iterations done (the callback count); returns 0 converged, maxiter not. */
static double dot_omp(long n, const double *a, const double *b)
{
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (long k = 0; k < n; ++k) s += a[k] * b[k];
    return s;
}

static void spmv_omp(long n, const int64_t *ptr, const int64_t *ind, const double *val,
                     const double *x, double *y)
{
#pragma omp parallel for schedule(static)
    for (long r = 0; r < n; ++r) {
        double s = 0.0;
        for (int64_t q = ptr[r]; q < ptr[r + 1]; ++q) s += val[q] * x[ind[q]];
        y[r] = s;
    }
}

int oracle_pcg_csr(long n, const int64_t *ptr, const int64_t *ind, const double *val,
                   const double *inv_diag, const double *b, double *x, double rtol, int maxiter,
                   double *work, int *iters)
{
    double *r = work, *z = work + n, *p = work + 2 * n, *q = work + 3 * n;
    spmv_omp(n, ptr, ind, val, x, q);
#pragma omp parallel for schedule(static)
    for (long k = 0; k < n; ++k) r[k] = b[k] - q[k];
    const double bnrm = sqrt(dot_omp(n, b, b));
    *iters = 0;
    if (bnrm == 0.0) {
        for (long k = 0; k < n; ++k) x[k] = b[k];
        return 0;
    }
    const double atol = rtol * bnrm;
    double rho_prev = 0.0;
    for (int it = 0; it < maxiter; ++it) {
        if (sqrt(dot_omp(n, r, r)) < atol) return 0;
#pragma omp parallel for schedule(static)
        for (long k = 0; k < n; ++k) z[k] = inv_diag[k] * r[k];
        const double rho = dot_omp(n, r, z);
        if (it > 0) {
            const double beta = rho / rho_prev;
#pragma omp parallel for schedule(static)
            for (long k = 0; k < n; ++k) p[k] = beta * p[k] + z[k];
        } else {
#pragma omp parallel for schedule(static)
            for (long k = 0; k < n; ++k) p[k] = z[k];
        }
        spmv_omp(n, ptr, ind, val, p, q);
        const double alpha = rho / dot_omp(n, p, q);
#pragma omp parallel for schedule(static)
        for (long k = 0; k < n; ++k) {
            x[k] += alpha * p[k];
            r[k] -= alpha * q[k];
        }
        rho_prev = rho;
        *iters = it + 1;
    }
    return maxiter;
}
