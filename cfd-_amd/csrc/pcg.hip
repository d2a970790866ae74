// Pressure projection as one persistent, cooperatively launched kernel:
//   right-hand side div(F, G)/dt with the Neumann compatibility shift
//   (solver.py:264-276, pressure.py:154-161), matrix-free Jacobi-PCG with the
//   iteration semantics of scipy.sparse.linalg.cg (iterative.py:305-423 in
//   SciPy 1.15, as called at pressure.py:171) or the FP32-bulk variant with
//   FP64 dots and FP64 true-residual replacement, the per-component pressure
//   mean (pressure.py:178-184), and the velocity correction with the
//   residual monitor (solver.py:277-321, 390-392).
//
// Convergence is decided on the device, so a whole solve costs one launch and
// no host round trip.  Each block owns a fixed band of grid rows for the whole
// launch (cache locality per XCD); phases are separated by the grid barrier of
// common.h, and every reduction is a fixed-order sum, so results are
// bit-reproducible run to run.
#include "cfd_kernels.h"

namespace {

struct Rows {
    int i0, i1;
};

__device__ __forceinline__ Rows block_rows(int nx) {
    Rows r;
    r.i0 = 1 + (int)(((long)nx * blockIdx.x) / gridDim.x);
    r.i1 = 1 + (int)(((long)nx * (blockIdx.x + 1)) / gridDim.x);
    return r;
}

// 5-point -Laplacian applied at cell (i, j); neighbours outside the fluid hold 0.
template <typename T>
__device__ __forceinline__ double apply_A(const T* __restrict__ d, int i, int j, int pc, int nx,
                                          int periodic, double diag, double ax, double ay) {
    const int iE = (periodic && i == nx) ? 1 : i + 1;
    const int iW = (periodic && i == 1) ? nx : i - 1;
    const double dE = (double)d[ix(iE, j, pc)], dW = (double)d[ix(iW, j, pc)];
    const double dN = (double)d[ix(i, j + 1, pc)], dS = (double)d[ix(i, j - 1, pc)];
    return diag * (double)d[ix(i, j, pc)] - ax * (dE + dW) - ay * (dN + dS);
}

__device__ __forceinline__ double sel_comp(const double* m, int c) {
    double v = 0.0;
#pragma unroll
    for (int k = 0; k < CFD_MAX_COMP; ++k) v = (c == k + 1) ? m[k] : v;
    return v;
}


// ---------------------------------------------------------------------------
// FP32 aggregation multigrid V-cycle (preconditioner of the FP64 flexible CG).
// Level 0 is the fine grid (matrix-free operator of the projection); level
// l >= 1 holds the Galerkin operator 0.5 P^T A P of 2x2-cell aggregation as a
// 5-point stencil (diag, east and north couplings); the coarsest level is
// solved with a dense inverse.  Smoother: damped Jacobi, two sweeps before
// and after the coarse correction (the first pre-sweep from x = 0 is local).
// Levels with many cells are spread over all blocks with grid syncs between
// sweeps; small levels run in block 0 alone with work-group barriers.
// ---------------------------------------------------------------------------
// Cells of a level handled by this block: a contiguous range of the
// row-major cell index (balanced over all blocks for grid levels, everything
// in block 0 otherwise); threads stride over the range.
struct MGPart {
    int k0, k1;
    bool active;
};

__device__ __forceinline__ MGPart mg_part(const MGLevelDev& L) {
    MGPart p;
    const long total = (long)L.nx * L.ny;
    if (L.grid) {
        p.k0 = (int)((total * blockIdx.x) / gridDim.x);
        p.k1 = (int)((total * (blockIdx.x + 1)) / gridDim.x);
        p.active = true;
    } else {
        p.k0 = 0;
        p.k1 = (int)total;
        p.active = blockIdx.x == 0;
    }
    return p;
}

#define MG_FOR(pt, L)                                                       \
    for (int k_ = (pt).k0 + (int)threadIdx.x; k_ < (pt).k1; k_ += blockDim.x) \
        for (int i = k_ / (L).ny + 1, j = k_ - (i - 1) * (L).ny + 1, once_ = 1; once_; once_ = 0)

struct MGCtx {
    const MGDev* mg;
    const uint8_t* code;     // fine-level codes (pitch of level 0)
    const float* diag0;      // fine diagonal table in LDS (64)
    int periodic;
    float ax, ay;
};

// diagonal of level l at cell c (0: not a fluid cell)
__device__ __forceinline__ float mg_diag(const MGCtx& m, int l, const MGLevelDev& L, long c) {
    return l == 0 ? m.diag0[m.code[c] & 63] : L.diag[c];
}

// (A v)(i, j) on level l; `val(i, j)` gives v at a cell (0 outside the fluid)
template <typename F>
__device__ __forceinline__ float mg_apply(const MGCtx& m, int l, const MGLevelDev& L, int i, int j,
                                          float d, F val) {
    const int iE = (m.periodic && i == L.nx) ? 1 : i + 1;
    const int iW = (m.periodic && i == 1) ? L.nx : i - 1;
    if (l == 0)
        return d * val(i, j) - m.ax * (val(iE, j) + val(iW, j)) - m.ay * (val(i, j + 1) + val(i, j - 1));
    const int P = L.P;
    return d * val(i, j) - L.cE[ix(i, j, P)] * val(iE, j) - L.cE[ix(iW, j, P)] * val(iW, j) -
           L.cN[ix(i, j, P)] * val(i, j + 1) - L.cN[ix(i, j - 1, P)] * val(i, j - 1);
}

__device__ __forceinline__ void mg_sync(const MGLevelDev& L, const GridSync& gs, double* lds,
                                        unsigned& epoch) {
    if (L.grid) grid_sync(gs, lds, epoch);
    else if (blockIdx.x == 0) __syncthreads();
}

// z = M r with the fine residual already in lev[0].b and lev[0].x = w b / d.
// Result in lev[0].t.
__device__ void mg_vcycle(const MGCtx& m, const GridSync& gs, double* lds, unsigned& epoch) {
    const MGDev& mg = *m.mg;
    const int NL = mg.nlev;
    const float w = mg.omega;
    const int tid = threadIdx.x, BS = blockDim.x;
    for (int l = 0; l < NL - 1; ++l) {
        const MGLevelDev& L = mg.lev[l];
        const MGPart pt = mg_part(L);
        mg_sync(L, gs, lds, epoch);
        if (pt.active) {   // second pre-smoothing sweep: t = x + w (b - A x) / d
            MG_FOR(pt, L) {
                    const long c = ix(i, j, L.P);
                    const float d = mg_diag(m, l, L, c);
                    if (d == 0.0f) continue;
                    auto xv = [&](int a, int b) -> float { return L.x[ix(a, b, L.P)]; };
                    const float Ax = mg_apply(m, l, L, i, j, d, xv);
                    L.t[c] = L.x[c] + w * (L.b[c] - Ax) / d;
                }
        }
        mg_sync(L, gs, lds, epoch);
        const MGLevelDev& C = mg.lev[l + 1];
        const MGPart pc = mg_part(C);
        if (pc.active) {   // restriction of the residual (sum over the 2x2 children)
            MG_FOR(pc, C) {
                    const int ci = i, cj = j;
                    const long cc = ix(ci, cj, C.P);
                    float bc = 0.0f;
                    for (int di = 0; di < 2; ++di)
                        for (int dj = 0; dj < 2; ++dj) {
                            const int i = 2 * ci - 1 + di, j = 2 * cj - 1 + dj;
                            if (i > L.nx || j > L.ny) continue;
                            const long c = ix(i, j, L.P);
                            const float d = mg_diag(m, l, L, c);
                            if (d == 0.0f) continue;
                            auto tv = [&](int a, int b) -> float { return L.t[ix(a, b, L.P)]; };
                            bc += L.b[c] - mg_apply(m, l, L, i, j, d, tv);
                        }
                    C.b[cc] = bc;
                    if (l + 1 < NL - 1) {
                        const float dc = C.diag[cc];
                        C.x[cc] = dc != 0.0f ? w * bc / dc : 0.0f;
                    }
                }
        }
    }
    // coarsest level: dense inverse (block 0)
    {
        const MGLevelDev& C = mg.lev[NL - 1];
        mg_sync(C, gs, lds, epoch);
        if (blockIdx.x == 0) {
            const int n = mg.ncoarse;
            for (int k = tid; k < n; k += BS) {
                float acc = 0.0f;
                for (int q = 0; q < n; ++q) {
                    const int qi = q / C.ny + 1, qj = q % C.ny + 1;
                    acc += mg.cinv[(long)k * n + q] * C.b[ix(qi, qj, C.P)];
                }
                C.t[ix(k / C.ny + 1, k % C.ny + 1, C.P)] = acc;
            }
        }
    }
    for (int l = NL - 2; l >= 0; --l) {
        const MGLevelDev& L = mg.lev[l];
        const MGLevelDev& C = mg.lev[l + 1];
        const MGPart pt = mg_part(L);
        mg_sync(L, gs, lds, epoch);
        // coarse correction t + wc e_parent, then the first post-smoothing sweep into x
        auto corr = [&](int a, int b) -> float {
            const long c = ix(a, b, L.P);
            if (mg_diag(m, l, L, c) == 0.0f) return 0.0f;
            return L.t[c] + mg.wc * C.t[ix((a + 1) >> 1, (b + 1) >> 1, C.P)];
        };
        if (pt.active) {
            MG_FOR(pt, L) {
                    const long c = ix(i, j, L.P);
                    const float d = mg_diag(m, l, L, c);
                    if (d == 0.0f) continue;
                    const float y = corr(i, j);
                    L.x[c] = y + w * (L.b[c] - mg_apply(m, l, L, i, j, d, corr)) / d;
                }
        }
        mg_sync(L, gs, lds, epoch);
        if (pt.active) {
            MG_FOR(pt, L) {
                    const long c = ix(i, j, L.P);
                    const float d = mg_diag(m, l, L, c);
                    if (d == 0.0f) continue;
                    auto xv = [&](int a, int b) -> float { return L.x[ix(a, b, L.P)]; };
                    L.t[c] = L.x[c] + w * (L.b[c] - mg_apply(m, l, L, i, j, d, xv)) / d;
                }
        }
    }
}

}  // namespace

__global__ void __launch_bounds__(PROJ_BS) k_project(ProjArgs a, Geo g) {
    __shared__ double lds[CFD_MAX_COMP * (PROJ_BS / 64) + 16];
    __shared__ double s_diag[64], s_inv[64];
    __shared__ float s_diagf[64];
    DevState* st = a.st;
    if (st->done) return;   // uniform over the grid: set by an earlier launch
    if (threadIdx.x < 64) {
        s_diag[threadIdx.x] = a.lut[threadIdx.x];
        s_inv[threadIdx.x] = a.lut[64 + threadIdx.x];
        s_diagf[threadIdx.x] = (float)a.lut[threadIdx.x];
    }
    __syncthreads();

    unsigned epoch = 0;
    const int nx = g.nx, ny = g.ny, pc = g.pc;
    const int periodic = g.periodic;
    const double ax = 1.0 / (g.dx * g.dx), ay = 1.0 / (g.dy * g.dy);
    const double dt = a.dt_override > 0.0 ? a.dt_override : st->dt;
    const Rows rows = block_rows(nx);
    const int tid = threadIdx.x, BS = blockDim.x;
    const bool lead = blockIdx.x == 0 && tid == 0;
    const GridSync gs{a.sync, a.sync + (long)gridDim.x * CFD_SLOT, a.spin};

#define FOR_CELLS                                            \
    for (int i = rows.i0; i < rows.i1; ++i)                  \
        for (int j = 1 + tid; j <= ny; j += BS)

    double cs[CFD_MAX_COMP];
    double mean[CFD_MAX_COMP];
    {
    // ---- right-hand side div(F,G)/dt and per-component sums ----
#pragma unroll
    for (int k = 0; k < CFD_MAX_COMP; ++k) cs[k] = 0.0;
    FOR_CELLS {
        const long c = ix(i, j, pc);
        if (!(a.code[c] & 63)) continue;
        double rhs;
        if (a.rhs_in) {
            rhs = a.rhs_in[c];
        } else {
            const double div = (a.F[ix(i, j, g.pu)] - a.F[ix(i - 1, j, g.pu)]) / g.dx +
                               (a.G[ix(i, j, g.pv)] - a.G[ix(i, j - 1, g.pv)]) / g.dy;
            rhs = div / dt;
        }
        a.b[c] = rhs;
        const int cm = a.comp[c];
#pragma unroll
        for (int k = 0; k < CFD_MAX_COMP; ++k) cs[k] += (cm == k + 1) ? rhs : 0.0;
    }
    if (a.ncomp > 0) {
        grid_sum<CFD_MAX_COMP>(cs, gs, lds, epoch);
#pragma unroll
        for (int k = 0; k < CFD_MAX_COMP; ++k)
            mean[k] = k < a.ncomp ? cs[k] / a.comp_count[k] : 0.0;
    } else {
#pragma unroll
        for (int k = 0; k < CFD_MAX_COMP; ++k) mean[k] = 0.0;
    }

    // ---- b = -(rhs - mean), r = b - A x, z = M r ----
    double s3[3] = {0.0, 0.0, 0.0};   // b.b, r.r, r.z
    FOR_CELLS {
        const long c = ix(i, j, pc);
        const int code = a.code[c];
        const int li = code & 63;
        if (!li) continue;
        double bv = a.b[c];
        const int cm = a.comp[c];
        if (cm) bv = bv - sel_comp(mean, cm);
        bv = -bv;
        a.b[c] = bv;
        const double Ax = apply_A(a.x, i, j, pc, nx, periodic, s_diag[li], ax, ay);
        const double rv = bv - Ax;
        a.r[c] = rv;
        const double z = s_inv[li] * rv;
        s3[0] += bv * bv;
        s3[1] += rv * rv;
        s3[2] += rv * z;
    }
    grid_sum<3>(s3, gs, lds, epoch);
    const double bnorm = sqrt(s3[0]);
    const double atol = a.rtol * bnorm;
    double rr = s3[1], rz = s3[2];
    int iters = 0, info = 0;

    if (bnorm == 0.0) {
        // SciPy returns b itself (all zeros) as the solution
        FOR_CELLS {
            const long c = ix(i, j, pc);
            if (a.code[c] & 63) a.x[c] = 0.0;
        }
        rr = 0.0;
    } else if (a.method == PROJ_PCG64) {
        bool conv = false;
        double rz_prev = 1.0;
        for (int k = 0; k < a.maxiter; ++k) {
            if (sqrt(rr) < atol || !(rr == rr)) {
                conv = sqrt(rr) < atol;
                break;
            }
            const double beta = k > 0 ? rz / rz_prev : 0.0;
            double* dnew = (k & 1) ? a.d1 : a.d0;
            const double* dold = (k & 1) ? a.d0 : a.d1;
            // direction update with the neighbours' new direction rebuilt from
            // r and the previous direction (no barrier between update and SpMV)
            double pq[1] = {0.0};
            FOR_CELLS {
                const long c = ix(i, j, pc);
                const int li = a.code[c] & 63;
                if (!li) continue;
                const int iE = (periodic && i == nx) ? 1 : i + 1;
                const int iW = (periodic && i == 1) ? nx : i - 1;
                const long cE = ix(iE, j, pc), cW = ix(iW, j, pc);
                const long cN = ix(i, j + 1, pc), cS = ix(i, j - 1, pc);
                auto dirv = [&](long cc) -> double {
                    const int l = a.code[cc] & 63;
                    const double z = s_inv[l] * a.r[cc];
                    return k > 0 ? z + beta * dold[cc] : z;
                };
                const double dc = dirv(c);
                const double q = s_diag[li] * dc - ax * (dirv(cE) + dirv(cW)) -
                                 ay * (dirv(cN) + dirv(cS));
                dnew[c] = dc;
                a.q[c] = q;
                pq[0] += dc * q;
            }
            grid_sum<1>(pq, gs, lds, epoch);
            const double alpha = rz / pq[0];
            double s2[2] = {0.0, 0.0};
            FOR_CELLS {
                const long c = ix(i, j, pc);
                const int li = a.code[c] & 63;
                if (!li) continue;
                const double dc = dnew[c];
                a.x[c] = a.x[c] + alpha * dc;
                const double rv = a.r[c] - alpha * a.q[c];
                a.r[c] = rv;
                const double z = s_inv[li] * rv;
                s2[0] += rv * rv;
                s2[1] += rv * z;
            }
            grid_sum<2>(s2, gs, lds, epoch);
            rz_prev = rz;
            rr = s2[0];
            rz = s2[1];
            iters = k + 1;
        }
        if (!conv && iters >= a.maxiter) info = a.maxiter;
        if (!conv && iters < a.maxiter) info = -1;   // breakdown (non-finite residual)
    } else {
        // flexible CG (Polak-Ribiere beta) with FP64 vectors and dot products,
        // preconditioned by the FP32 multigrid V-cycle; the loop stops on the
        // recursive residual, then the true residual b - A x decides (a restart
        // with the true residual if it is not yet below the tolerance)
        const MGCtx m{a.mg, a.code, s_diagf, periodic, (float)ax, (float)ay};
        const MGLevelDev& L0 = a.mg->lev[0];
        const float w = a.mg->omega;
        double rho_old = 1.0;
        bool conv = false;
        int k = 0, kstart = 0, reps = 0;
        for (int rep = 0; rep < 3; ++rep) {
            for (; k < a.maxiter; ++k) {
                if (sqrt(rr) < atol || !(rr == rr)) break;
                FOR_CELLS {
                    const long c = ix(i, j, pc);
                    const int li = a.code[c] & 63;
                    if (!li) continue;
                    const float rf = (float)a.r[c];
                    L0.b[c] = rf;
                    L0.x[c] = w * rf / s_diagf[li];
                }
                mg_vcycle(m, gs, lds, epoch);
                const bool first = k == kstart;
                double s2[2] = {0.0, 0.0};
                FOR_CELLS {
                    const long c = ix(i, j, pc);
                    if (!(a.code[c] & 63)) continue;
                    const double z = (double)L0.t[c];
                    s2[0] += a.r[c] * z;
                    if (!first) s2[1] += a.rold[c] * z;
                }
                grid_sum<2>(s2, gs, lds, epoch);
                const double beta = first ? 0.0 : (s2[0] - s2[1]) / rho_old;
                double* dnew = (k & 1) ? a.d1 : a.d0;
                const double* dold = (k & 1) ? a.d0 : a.d1;
                double pq[1] = {0.0};
                FOR_CELLS {
                    const long c = ix(i, j, pc);
                    const int li = a.code[c] & 63;
                    if (!li) continue;
                    const int iE = (periodic && i == nx) ? 1 : i + 1;
                    const int iW = (periodic && i == 1) ? nx : i - 1;
                    auto dirv = [&](long cc) -> double {
                        const double z = (double)L0.t[cc];
                        return first ? z : z + beta * dold[cc];
                    };
                    const double dc = dirv(c);
                    const double q = s_diag[li] * dc - ax * (dirv(ix(iE, j, pc)) + dirv(ix(iW, j, pc))) -
                                     ay * (dirv(ix(i, j + 1, pc)) + dirv(ix(i, j - 1, pc)));
                    dnew[c] = dc;
                    a.q[c] = q;
                    pq[0] += dc * q;
                }
                grid_sum<1>(pq, gs, lds, epoch);
                const double alpha = s2[0] / pq[0];
                double r2[1] = {0.0};
                FOR_CELLS {
                    const long c = ix(i, j, pc);
                    if (!(a.code[c] & 63)) continue;
                    a.x[c] = a.x[c] + alpha * dnew[c];
                    const double rv = a.r[c];
                    a.rold[c] = rv;
                    const double rn = rv - alpha * a.q[c];
                    a.r[c] = rn;
                    r2[0] += rn * rn;
                }
                grid_sum<1>(r2, gs, lds, epoch);
                rr = r2[0];
                rho_old = s2[0];
            }
            // true residual of the FP64 solution
            grid_sync(gs, lds, epoch);
            double t2[1] = {0.0};
            FOR_CELLS {
                const long c = ix(i, j, pc);
                const int li = a.code[c] & 63;
                if (!li) continue;
                const double rt = a.b[c] - apply_A(a.x, i, j, pc, nx, periodic, s_diag[li], ax, ay);
                a.r[c] = rt;
                t2[0] += rt * rt;
            }
            grid_sum<1>(t2, gs, lds, epoch);
            rr = t2[0];
            reps += 1;
            conv = sqrt(rr) < atol;
            if (conv || k >= a.maxiter || !(rr == rr)) break;
            kstart = k;
        }
        iters = k;
        info = conv ? 0 : (k >= a.maxiter ? a.maxiter : -1);
        if (lead) st->cg_outer = reps;
    }
    if (lead) {
        st->cg_iters = iters;
        st->cg_info = info;
        if (a.method == PROJ_PCG64) st->cg_outer = 0;
        st->cg_bnorm = bnorm;
        st->cg_relres = bnorm > 0.0 ? sqrt(rr) / bnorm : 0.0;
    }
    }

    // ---- pressure = x minus its mean per Neumann component ----
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
    // rows of this block (ghost columns 0 and ny+1 are zero), plus the ghost
    // rows 0 and nx+1 by the first and last block (periodic copies of cells)
    for (int i = rows.i0; i < rows.i1; ++i)
        for (int j = tid; j <= ny + 1; j += BS)
            a.p[ix(i, j, pc)] = (j >= 1 && j <= ny) ? pval(i, j) : 0.0;
    if (blockIdx.x == 0)
        for (int j = tid; j <= ny + 1; j += BS)
            a.p[ix(0, j, pc)] = (periodic && j >= 1 && j <= ny) ? pval(nx, j) : 0.0;
    if (blockIdx.x == gridDim.x - 1)
        for (int j = tid; j <= ny + 1; j += BS)
            a.p[ix(nx + 1, j, pc)] = (periodic && j >= 1 && j <= ny) ? pval(1, j) : 0.0;
    if (!a.correct) return;
    grid_sync(gs, lds, epoch);

    // ---- velocity correction u = F - dt grad p and the residual monitor ----
    const double* p = a.p;
    double res[1] = {0.0};
    auto fluid = [&](int i, int j) -> bool { return a.code[ix(i, j, pc)] & 63; };
    for (int i = rows.i0; i < rows.i1; ++i) {
        for (int j = 1 + tid; j <= ny; j += BS) {
            // u face (i, j): interior faces, the periodic face nx, or the right boundary
            const long cu = ix(i, j, g.pu);
            const double F = a.F[cu];
            double un;
            bool uf = false;
            if (i < nx) {
                uf = fluid(i, j) && fluid(i + 1, j);
                un = uf ? F - dt * (p[ix(i + 1, j, pc)] - p[ix(i, j, pc)]) / g.dx : F;
            } else if (periodic) {
                uf = fluid(nx, j) && fluid(1, j);
                un = uf ? F - dt * (p[ix(1, j, pc)] - p[ix(nx, j, pc)]) / g.dx : F;
            } else {
                un = (a.code[ix(nx, j, pc)] & 64) ? F + 2 * dt * p[ix(nx, j, pc)] / g.dx : F;
            }
            if (uf) res[0] = nanmax(res[0], fabs(un - a.u[cu]));
            a.u[cu] = un;
            if (periodic && i == nx) a.u[ix(0, j, g.pu)] = un;
        }
        // v faces (i, j), j in 0..ny
        for (int j = tid; j <= ny; j += BS) {
            const long cv = ix(i, j, g.pv);
            const double G = a.G[cv];
            double vn;
            bool vf = false;
            if (j >= 1 && j < ny) {
                vf = fluid(i, j) && fluid(i, j + 1);
                vn = vf ? G - dt * (p[ix(i, j + 1, pc)] - p[ix(i, j, pc)]) / g.dy : G;
            } else if (periodic) {
                continue;   // solver.py:300-306 returns before the boundary faces
            } else if (j == 0) {
                vn = (a.code[ix(i, 1, pc)] & 128) ? G - 2 * dt * p[ix(i, 1, pc)] / g.dy : G;
            } else {
                vn = (a.code[ix(i, ny, pc)] & 128) ? G + 2 * dt * p[ix(i, ny, pc)] / g.dy : G;
            }
            if (vf) res[0] = nanmax(res[0], fabs(vn - a.v[cv]));
            a.v[cv] = vn;
        }
    }
    if (blockIdx.x == 0) {
        if (!periodic) {
            for (int j = 1 + tid; j <= ny; j += BS) {
                const long cu = ix(0, j, g.pu);
                const double F = a.F[cu];
                a.u[cu] = (a.code[ix(1, j, pc)] & 64) ? F - 2 * dt * p[ix(1, j, pc)] / g.dx : F;
            }
        }
    }
    grid_max<1>(res, gs, lds, epoch);
    if (lead && a.dt_override <= 0.0) st->residual = res[0];
    if (periodic && blockIdx.x == gridDim.x - 1 && tid < 2) {
        const int j = tid == 0 ? 0 : ny + 1;
        a.u[ix(0, j, g.pu)] = a.u[ix(nx, j, g.pu)];
    }
    if (a.zero_p_after) {
        for (int i = rows.i0; i < rows.i1; ++i)
            for (int j = tid; j <= ny + 1; j += BS) a.p[ix(i, j, pc)] = 0.0;
        if (blockIdx.x == 0)
            for (int j = tid; j <= ny + 1; j += BS) a.p[ix(0, j, pc)] = 0.0;
        if (blockIdx.x == gridDim.x - 1)
            for (int j = tid; j <= ny + 1; j += BS) a.p[ix(nx + 1, j, pc)] = 0.0;
    }
#undef FOR_CELLS
}
