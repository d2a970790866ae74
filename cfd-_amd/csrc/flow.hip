// Flow-rate control of periodic channels (cfd2d/solver.py:405-425, the
// OpenFOAM meanVelocityForce idea): after the pressure correction the
// velocity on the fluid faces is shifted by du = U_target - U_bulk, du / dt
// is added to the driving force of the next steps, and with obstacles the
// shifted field is projected again with dt = 1, keeping p.  Because the
// shift changes u after the projection, the step's residual
// max |u_new - u_start| / dt (solver.py:390-392) is taken here, against a
// copy of the velocity from the start of the step.
#pragma clang fp contract(off)
#include "cfd_kernels.h"

struct FlowArgs {
    double* u;
    const double* v;
    const uint8_t* fl;
    DevState* st;
    double flow_rate;
    double nfaces;           // fluid u faces 1..nx (the bulk velocity's count)
    double* part;            // per-block partial sums
    unsigned* ticket;        // [0] bulk sum, [1] residual
    unsigned long long* acc; // residual maximum (bits of a non-negative double)
    const double* u0;        // velocity at the start of the step
    const double* v0;
};

__device__ __forceinline__ bool u_fluid_face(const uint8_t* fl, const Geo& g, int i, int j) {
    return i >= 1 && (fl[ix(i, j, g.pc)] & 1) && (fl[ix(i + 1, j, g.pc)] & 1);
}

// bulk velocity: the mean of u over the fluid faces i = 1..nx, j = 1..ny
// (solver.py:405-408); partial sums per block, added in block order by the
// last block, which sets du and the driving force
__global__ void __launch_bounds__(256) k_flow_bulk(FlowArgs a, Geo g) {
    __shared__ double lds[16];
    if (a.st->done) return;
    const int nx = g.nx, ny = g.ny;
    const int r0 = 1 + (int)(((long)nx * blockIdx.x) / gridDim.x);
    const int r1 = 1 + (int)(((long)nx * (blockIdx.x + 1)) / gridDim.x);
    double s[1] = {0.0};
    for (int i = r0; i < r1; ++i)
        for (int j = 1 + threadIdx.x; j <= ny; j += blockDim.x)
            if (u_fluid_face(a.fl, g, i, j)) s[0] += a.u[ix(i, j, g.pu)];
    block_sum<1>(s, lds);
    __shared__ int s_last;
    if (threadIdx.x == 0) {
        a.part[blockIdx.x] = s[0];
        const unsigned tk =
            __hip_atomic_fetch_add(&a.ticket[0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        s_last = tk == gridDim.x - 1;
    }
    __syncthreads();
    if (!s_last) return;
    double t[1] = {0.0};
    for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x)
        t[0] += __hip_atomic_load(&a.part[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    block_sum<1>(t, lds);
    if (threadIdx.x != 0) return;
    a.ticket[0] = 0u;
    const double du = a.flow_rate - t[0] / a.nfaces;
    a.st->flow_du = du;
    a.st->flow_force += du / a.st->dt;
}

// u += du on the fluid faces; periodic: face 0 repeats face nx
__global__ void __launch_bounds__(256) k_flow_shift(FlowArgs a, Geo g) {
    const int2 rc_ = xcd_rowgrid();
    const int j = rc_.x * blockDim.x + threadIdx.x;
    const int i = rc_.y;
    if (j > g.ny + 1 || i > g.nx) return;
    if (a.st->done) return;
    const int c = ix(i, j, g.pu);
    double x = a.u[c];
    if (u_fluid_face(a.fl, g, i, j)) {
        x = x + a.st->flow_du;
        a.u[c] = x;
    }
    if (g.periodic && i == g.nx) a.u[ix(0, j, g.pu)] = x;
}

// residual: max |u - u0| over the fluid u faces and |v - v0| over the fluid
// v faces (solver.py:138-142, 390-391); face 0 is not a fluid face when
// periodic (it repeats face nx)
__global__ void __launch_bounds__(256) k_flow_residual(FlowArgs a, Geo g) {
    __shared__ double lds[2 * 16];
    if (a.st->done) return;
    const int nx = g.nx, ny = g.ny;
    const int r0 = (int)(((long)(nx + 2) * blockIdx.x) / gridDim.x);
    const int r1 = (int)(((long)(nx + 2) * (blockIdx.x + 1)) / gridDim.x);
    double m[2] = {0.0, 0.0};
    for (int i = r0; i < r1; ++i)
        for (int j = threadIdx.x; j <= ny + 1; j += blockDim.x) {
            if (i <= nx && u_fluid_face(a.fl, g, i, j))
                m[0] = nanmax(m[0], fabs(a.u[ix(i, j, g.pu)] - a.u0[ix(i, j, g.pu)]));
            if (j <= ny && (a.fl[ix(i, j, g.pc)] & 1) && (a.fl[ix(i, j + 1, g.pc)] & 1))
                m[1] = nanmax(m[1], fabs(a.v[ix(i, j, g.pv)] - a.v0[ix(i, j, g.pv)]));
        }
    m[0] = nanmax(m[0], m[1]);
    double mm[1] = {m[0]};
    block_max<1>(mm, lds);
    __shared__ int s_last;
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_max(a.acc, (unsigned long long)__double_as_longlong(mm[0]),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned tk =
            __hip_atomic_fetch_add(&a.ticket[1], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        s_last = tk == gridDim.x - 1;
    }
    __syncthreads();
    if (!s_last || threadIdx.x != 0) return;
    const double r = __longlong_as_double(
        (long long)__hip_atomic_load(a.acc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    *a.acc = 0ull;
    a.ticket[1] = 0u;
    a.st->residual = r;
}
