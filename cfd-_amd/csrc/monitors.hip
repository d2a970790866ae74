// Function objects evaluated on the device after every step
// (cfd2d/functions.py:59-240 evaluates them on the host from the fields):
// point probes of u, v, the cell pressure, speed, vorticity and scalars by
// bilinear interpolation (functions.py:24-39, 59-114: scipy's linear
// RegularGridInterpolator), and the time-weighted field sums of FieldAverage
// (functions.py:180-210).  They run in the step's launch sequence, so a run
// with them needs no per-step host synchronisation; the host drains the
// probe records when it synchronises anyway.
#include "cfd_kernels.h"

constexpr int MON_MAX_SCALARS = CFD_MAX_SCALARS;

// field codes of a probe sample (CFD_PROBE_* of the C interface)
enum : int { PR_U = 0, PR_V = 1, PR_P = 2, PR_SPEED = 3, PR_VORT = 4, PR_SCALAR = 5 };

struct ProbeArgs {
    const double* u;
    const double* v;
    const double* p;
    const uint8_t* fl;
    const double* phi[MON_MAX_SCALARS];
    const DevState* st;
    const long long* last_n;   // step count at the previous monitor pass
    int nsample, every, cap, P;
    double dx, dy;
    const int* src;            // per sample: PR_*
    const long long* base;     // per sample: device offset of the lower corner
    const double* frac;        // per sample (a, b)
    double* times;             // [cap]
    double* vals;              // [cap][nsample]
    int* count;                // records held; lost records counted in count[1]
};

__device__ __forceinline__ double qnan() { return __longlong_as_double(0x7ff8000000000000ll); }

// vorticity dv/dx - du/dy at the grid node whose u face below it and v face
// left of it sit at device offset o (solver.py:569-573)
__device__ __forceinline__ double node_vort(const ProbeArgs& a, long long o) {
    return (a.v[o + a.P] - a.v[o]) / a.dx - (a.u[o + 1] - a.u[o]) / a.dy;
}

// a cell-centred field at the cell with device offset c, NaN in solid cells
// (solver.py:526-537, 575-578)
__device__ double cell_value(const ProbeArgs& a, int sr, long long c) {
    if (!(a.fl[c] & 1)) return qnan();
    if (sr == PR_P) return a.p[c];
    if (sr == PR_SPEED) {
        const double uc = 0.5 * (a.u[c - a.P] + a.u[c]);
        const double vc = 0.5 * (a.v[c - 1] + a.v[c]);
        return sqrt(uc * uc + vc * vc);
    }
    if (sr == PR_VORT) {
        const long long o = c - a.P - 1;
        return 0.25 * (node_vort(a, o) + node_vort(a, o + a.P) + node_vort(a, o + 1) +
                       node_vort(a, o + a.P + 1));
    }
    return a.phi[sr - PR_SCALAR][c];
}

__device__ __forceinline__ double probe_value(const ProbeArgs& a, int sr, long long o) {
    if (sr == PR_U) return a.u[o];
    if (sr == PR_V) return a.v[o];
    return cell_value(a, sr, o);
}

// one work-group: after a step that ran (st->n moved) with n % every == 0,
// one record: t and the interpolated values, in the reference's operation
// order f00 (1-a)(1-b) + f01 (1-a) b + f10 a (1-b) + f11 a b (a NaN corner
// makes the sample NaN, whatever its weight, as it does there)
__global__ void __launch_bounds__(256) k_probes(ProbeArgs a) {
    __shared__ int s_slot;
    if (threadIdx.x == 0) {
        const long long n = a.st->n;
        int slot = -1;
        if (n != *a.last_n && n % a.every == 0) {
            if (a.count[0] < a.cap) {
                slot = a.count[0];
                a.count[0] = slot + 1;
                a.times[slot] = a.st->t;
            } else {
                a.count[1] += 1;
            }
        }
        s_slot = slot;
    }
    __syncthreads();
    const int slot = s_slot;
    if (slot < 0) return;
    for (int s = threadIdx.x; s < a.nsample; s += blockDim.x) {
        const int sr = a.src[s];
        const long long o = a.base[s];
        const double fa = a.frac[2 * s], fb = a.frac[2 * s + 1];
        const double f00 = probe_value(a, sr, o), f01 = probe_value(a, sr, o + 1);
        const double f10 = probe_value(a, sr, o + a.P), f11 = probe_value(a, sr, o + a.P + 1);
        const double ca = 1.0 - fa, cb = 1.0 - fb;
        const double val = f00 * ca * cb + f01 * ca * fb + f10 * fa * cb + f11 * fa * fb;
        a.vals[(long)slot * a.nsample + s] = val;
    }
}

struct FavgArgs {
    const double* u;
    const double* v;
    const double* p;
    const uint8_t* fl;
    const DevState* st;
    const long long* last_n;
    double t_start;
    double* sum;     // [3][nx][ny]: u, v, p at the cell centres (host layout)
    double* sum2;
    double* T;       // accumulated time
};

// after a step that ran with t >= t_start: sum += a dt, sum2 += a a dt for
// the cell-centred u, v and pressure (NaN in solid cells); grid (nx, ny)
// cells, one per thread
__global__ void __launch_bounds__(256) k_favg(FavgArgs a, Geo g) {
    const DevState* st = a.st;
    if (st->n == *a.last_n || st->t < a.t_start) return;
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long ncell = (long)g.nx * g.ny;
    const double dt = st->dt;
    if (k == 0) *a.T += dt;
    if (k >= ncell) return;
    const int i = (int)(k / g.ny) + 1, j = (int)(k % g.ny) + 1;
    const bool solid = !(a.fl[ix(i, j, g.pc)] & 1);
    const double nan = qnan();
    double val[3];
    val[0] = 0.5 * (a.u[ix(i - 1, j, g.pu)] + a.u[ix(i, j, g.pu)]);
    val[1] = 0.5 * (a.v[ix(i, j - 1, g.pv)] + a.v[ix(i, j, g.pv)]);
    val[2] = a.p[ix(i, j, g.pc)];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        const double x = solid ? nan : val[f];
        a.sum[f * ncell + k] += x * dt;
        a.sum2[f * ncell + k] += x * x * dt;
    }
}

// the step count the monitors compare against, after all monitor kernels
__global__ void k_mon_fin(const DevState* st, long long* last_n) { *last_n = st->n; }

// ---- fields for reports, computed where the velocity lives ----------------
// Node vorticity dv/dx - du/dy (solver.py:569-573) into a dense
// (nx+1) x (ny+1) array; node (i, j) sits between u faces (i, j), (i, j+1)
// and v faces (i, j), (i+1, j) of the device layout.
__global__ void __launch_bounds__(256) k_vort_nodes(const double* __restrict__ u,
                                                    const double* __restrict__ v, Geo g, int P,
                                                    double* __restrict__ out) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int cols = g.ny + 1;
    if (k >= (long)(g.nx + 1) * cols) return;
    const long o = (k / cols) * P + k % cols;
    out[k] = (v[o + P] - v[o]) / g.dx - (u[o + 1] - u[o]) / g.dy;
}

// Streamfunction with u = dpsi/dy, v = -dpsi/dx and psi(0, 0) = 0
// (solver.py:580-596): faces with no fluid cell beside them carry no flux.
// Bottom row first: psi(i+1, 0) = -(sum_{k<=i} v0_k) dx, the sum running
// left to right as numpy's cumsum does; the work-group stages the open
// bottom faces 1024 at a time in LDS and one lane carries the sum.
__global__ void __launch_bounds__(1024) k_psi_bottom(const double* __restrict__ v,
                                                     const uint8_t* __restrict__ fl, Geo g, int P,
                                                     double* __restrict__ psi) {
    __shared__ double s_v[1024];
    const int cols = g.ny + 1;
    double c = 0.0;
    if (threadIdx.x == 0) psi[0] = 0.0;
    for (int base = 0; base < g.nx; base += 1024) {
        const int i = base + (int)threadIdx.x;
        if (i < g.nx) {
            const long o = (long)(i + 1) * P;
            s_v[threadIdx.x] = (fl[o] | fl[o + 1]) ? v[o] : 0.0;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const int m = min(1024, g.nx - base);
            for (int k = 0; k < m; ++k) {
                c = (base + k == 0) ? s_v[0] : c + s_v[k];
                psi[(long)(base + k + 1) * cols] = -c * g.dx;
            }
        }
        __syncthreads();
    }
}

// then every column upwards from its bottom value: psi(i, j+1) =
// psi(i, 0) + (sum_{k<=j} u_ik) dy, one lane per column
__global__ void __launch_bounds__(256) k_psi_cols(const double* __restrict__ u,
                                                  const uint8_t* __restrict__ fl, Geo g, int P,
                                                  double* __restrict__ psi) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > g.nx) return;
    const int cols = g.ny + 1;
    const long o = (long)i * P;
    double* col = psi + (long)i * cols;
    const double p0 = col[0];
    double s = 0.0;
    for (int j = 0; j < g.ny; ++j) {
        const bool open = fl[o + j + 1] | fl[o + P + j + 1];
        const double x = open ? u[o + j + 1] : 0.0;
        s = j == 0 ? x : s + x;
        col[j + 1] = p0 + s * g.dy;
    }
}
