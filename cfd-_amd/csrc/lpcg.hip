// Learned flexible CG for the pressure system (pressure_solver "cg_learned").
//
// Outer solver: flexible CG (Polak-Ribiere beta) with FP64 vectors and dot
// products, ending on the FP64 true residual ||b - A x|| < rtol ||b|| (with a
// restart from the true residual otherwise), as in the multigrid-preconditioned
// branch of k_project.  Its iterations are ordinary kernel launches, captured
// once into a HIP graph and replayed; after the solve has finished every
// kernel returns at once (the LDS coarse end only writes nothing back).
// Reductions: each work-group stores its partial sums,
// and the kernel that needs the total adds them in work-group order (so the
// results are reproducible); one-work-group kernels take the decisions that
// the host reads (converged, restart, breakdown).
//
// Preconditioner: an FP32 V-cycle over the aggregation hierarchy (damped
// Jacobi, two sweeps down and up, as mg_vcycle) that stops at a coarse level K,
// where the coarse correction is
//     e_K = c 1 + (sigma / D) s NN((b_K - c A_K 1) / s)
// followed by two Jacobi sweeps on level K.  c = 1^T b_K / 1^T A_K 1 is the
// Galerkin correction in the constant vector (the near-null space of the
// pinned Neumann problem), s = RMS of the deflated right-hand side, D the
// interior diagonal 2 (1/dx^2 + 1/dy^2), sigma the network's output scale, and
// NN the U-Net of unet.hip (an image of the level-K cells padded to multiples
// of 8).  Without the network (mode LEARNED_NONE) the coarse correction is
// c 1 alone; mode LEARNED_VCYCLE continues the V-cycle below level K to the
// dense coarsest level instead (the complete multigrid preconditioner of
// cg_mixed, the method the network has to beat); the teacher mode (LEARNED_TEACHER) solves level K to high accuracy
// (flexible CG preconditioned by V-cycles down to the dense coarsest level)
// and can store the pair (network input, normalised correction) for training.
#include "cfd_kernels.h"

enum { LEARNED_NET = 0, LEARNED_NONE = 1, LEARNED_TEACHER = 2, LEARNED_VCYCLE = 3 };

// iteration after which a network that has not reduced ||r|| tenfold is
// switched off (the solver then runs as without it)
constexpr int LV_NET_CHECK = 8;

// partial-sum slots (LV_SLOT doubles each) of LvArgs::part
enum { LV_P_DOTS = 0, LV_P_PQ = 1, LV_P_RR = 2, LV_P_CM = 3, LV_NSLOT = 4 };
constexpr int LV_MAXB = 1024;            // work-groups of any reducing kernel
constexpr int LV_SLOT = LV_MAXB * 8;

struct LvArgs {
    const MGDev* mg;         // hierarchy (device memory)
    const uint8_t* code;     // fine-level codes
    const float* diag0;      // fine diagonal table (64 floats)
    int periodic;
    float ax, ay;
    FcgState* fs;
    double* part;            // LV_NSLOT slots of partial sums
    int store_x0;            // the next V-cycle's level-0 x = w b / d is read (launch
                             // sequence); the fused down-leg forms it from b itself
};

struct FcgVec {
    double *x, *b, *r, *rold, *q, *d0, *d1;
    const double* lut;       // fine diagonal table (double)
    int nx, pc;
    double ax, ay;
};

namespace {

__device__ __forceinline__ MGCtx lv_ctx(const LvArgs& a) {
    return MGCtx{a.mg, a.code, a.diag0, a.periodic, a.ax, a.ay};
}

// the fine diagonal table in LDS (after the uniform early return)
#define LV_CTX_LDS(a)                                                          \
    __shared__ float s_d0_[64];                                                \
    if (threadIdx.x < 64) s_d0_[threadIdx.x] = (a).diag0[threadIdx.x];         \
    __syncthreads();                                                           \
    const MGCtx m{(a).mg, (a).code, s_d0_, (a).periodic, (a).ax, (a).ay}

// f(i, j) over the level's interior cells: a band of rows per work-group,
// threads along the contiguous j direction
template <typename F>
__device__ __forceinline__ void lv_cells(int nx, int ny, F f) {
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int i0 = 1 + (int)(((long)nx * b) / gridDim.x);
    const int i1 = 1 + (int)(((long)nx * (b + 1)) / gridDim.x);
    for (int i = i0; i < i1; ++i)
        for (int j = 1 + (int)threadIdx.x; j <= ny; j += blockDim.x) f(i, j);
}

// The same band of cells, U cells per thread at a time: gather(i, j) (loads
// and arithmetic only, no stores, no branches) runs for all U cells before
// scatter(i, j, value) stores any of them, so the loads of U cells are in
// flight together instead of one cell's chain of dependent loads after the
// other.  A thread visits its cells in the order of lv_cells when ny is a
// multiple of the work-group size.
template <int U, typename G, typename S>
__device__ __forceinline__ void lv_cells_batch(int nx, int ny, G gather, S scatter) {
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int i0 = 1 + (int)(((long)nx * b) / gridDim.x);
    const int i1 = 1 + (int)(((long)nx * (b + 1)) / gridDim.x);
    const int ncell = (i1 - i0) * ny, bs = blockDim.x;
    using T = decltype(gather(1, 1));
    for (int q0 = threadIdx.x; q0 < ncell; q0 += U * bs) {
        T val[U];
        int ii[U], jj[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int q = q0 + u * bs < ncell ? q0 + u * bs : q0;   // a valid cell
            ii[u] = i0 + q / ny;
            jj[u] = 1 + q % ny;
            val[u] = gather(ii[u], jj[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (q0 + u * bs < ncell) scatter(ii[u], jj[u], val[u]);
    }
}

// f(i, j) for one cell per thread: work-group = (row, 256-cell chunk of the
// row), launched with nx * ceil(ny / 256) work-groups of 256 threads
template <typename F>
__device__ __forceinline__ void lv_cell1(int nx, int ny, F f) {
    const int cj = (ny + 255) >> 8;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int i = b / cj + 1;
    const int j = (b - (i - 1) * cj) * 256 + (int)threadIdx.x + 1;
    if (i <= nx && j <= ny) f(i, j);
}

// this work-group's sums of v into slot[blockIdx.x * NV + q]
template <int NV>
__device__ __forceinline__ void lv_partial(double (&v)[NV], double* slot) {
    __shared__ double lds[NV * 16];
    block_sum<NV>(v, lds);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < NV; ++q) slot[(long)blockIdx.x * NV + q] = v[q];
}

// totals of the partials of nb work-groups (every thread gets them)
template <int NV>
__device__ __forceinline__ void lv_total(const double* slot, int nb, double (&v)[NV]) {
    __shared__ double lds[NV * 16];
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
#pragma unroll
        for (int q = 0; q < NV; ++q) v[q] += slot[(long)b * NV + q];
    block_sum<NV>(v, lds);
}

// Block sums of v written to part; the last work-group to arrive returns true
// with the totals (summed in block order) in v.  (Teacher solve only.)
template <int NV>
__device__ bool lv_reduce(double (&v)[NV], double* part, unsigned* ticket) {
    __shared__ double lds[NV * 16];
    __shared__ int s_last;
    block_sum<NV>(v, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) part[(long)blockIdx.x * NV + q] = v[q];
        __threadfence();
        const unsigned t =
            __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        s_last = t == gridDim.x - 1;
    }
    __syncthreads();
    if (!s_last) return false;
    __threadfence();
    double acc[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x)
#pragma unroll
        for (int q = 0; q < NV; ++q) acc[q] += part[(long)b * NV + q];
    block_sum<NV>(acc, lds);
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = acc[q];
    if (threadIdx.x == 0) *ticket = 0u;
    return true;
}

// (A 1)(i, j) on level l: the row sum over fluid neighbours
__device__ __forceinline__ float lv_a1(const MGCtx& m, int l, const MGLevelDev& L, int i, int j,
                                       float d) {
    auto one = [&](int a, int b) -> float {
        return mg_diag(m, l, L, ix(a, b, L.P)) != 0.0f ? 1.0f : 0.0f;
    };
    return mg_apply(m, l, L, i, j, d, one);
}

}  // namespace

// dst = src + w (b - A src) / d on level l (from_t: src = t, dst = x; else x -> t)
__global__ void __launch_bounds__(256) k_lv_smooth(LvArgs a, int l, int from_t) {
    if (a.fs->done) return;
    LV_CTX_LDS(a);
    const MGLevelDev L = a.mg->lev[l];
    const float w = a.mg->omega;
    const float* src = from_t ? L.t : L.x;
    float* dst = from_t ? L.x : L.t;
    lv_cell1(L.nx, L.ny, [&](int i, int j) {
        const long c = ix(i, j, L.P);
        const float d = mg_diag(m, l, L, c);
        if (d == 0.0f) return;
        auto sv = [&](int p, int q) -> float { return src[ix(p, q, L.P)]; };
        dst[c] = src[c] + w * (L.b[c] - mg_apply(m, l, L, i, j, d, sv)) / d;
    });
}

// last sweep of the V-cycle on the fine level, z = t = x + w (b - A x) / d,
// with the partial dot products r.z and r_old.z of the flexible CG
__global__ void __launch_bounds__(512) k_lv_smooth_dots(LvArgs a, FcgVec v) {
    const FcgState* fs = a.fs;
    if (fs->done) return;
    LV_CTX_LDS(a);
    const MGLevelDev L = a.mg->lev[0];
    const float w = a.mg->omega;
    const bool first = fs->first;
    double s[2] = {0.0, 0.0};
    struct Z {
        float d, z;
        double r, ro;
    };
    lv_cells_batch<4>(
        L.nx, L.ny,
        [&](int i, int j) {
            const long c = ix(i, j, L.P);
            Z o;
            o.d = mg_diag(m, 0, L, c);
            auto xv = [&](int p, int q) -> float { return L.x[ix(p, q, L.P)]; };
            o.z = L.x[c] + w * (L.b[c] - mg_apply(m, 0, L, i, j, o.d, xv)) / o.d;
            o.r = v.r[c];
            o.ro = v.rold[c];
            return o;
        },
        [&](int i, int j, const Z& o) {
            if (o.d == 0.0f) return;
            L.t[ix(i, j, L.P)] = o.z;
            s[0] += o.r * (double)o.z;
            if (!first) s[1] += o.ro * (double)o.z;
        });
    lv_partial<2>(s, a.part + LV_P_DOTS * LV_SLOT);
}

// partial dot products r.z and r_old.z with z = lev[0].t already computed
__global__ void __launch_bounds__(512) k_fcg_dots(LvArgs a, FcgVec v) {
    const FcgState* fs = a.fs;
    if (fs->done) return;
    const float* z = a.mg->lev[0].t;
    const bool first = fs->first;
    double s[2] = {0.0, 0.0};
    lv_cells(v.nx, a.mg->lev[0].ny, [&](int i, int j) {
        const long c = ix(i, j, v.pc);
        if (!(a.code[c] & 63)) return;
        const double zc = z[c];
        s[0] += v.r[c] * zc;
        if (!first) s[1] += v.rold[c] * zc;
    });
    lv_partial<2>(s, a.part + LV_P_DOTS * LV_SLOT);
}

// level l+1: b = sum over the 2x2 children of (b - A t); x = w b / d
__global__ void __launch_bounds__(256) k_lv_restrict(LvArgs a, int l) {
    if (a.fs->done) return;
    LV_CTX_LDS(a);
    const MGLevelDev L = a.mg->lev[l];
    const MGLevelDev C = a.mg->lev[l + 1];
    const float w = a.mg->omega;
    lv_cell1(C.nx, C.ny, [&](int ci, int cj) {
        const long cc = ix(ci, cj, C.P);
        float bc = 0.0f;
        for (int di = 0; di < 2; ++di)
            for (int dj = 0; dj < 2; ++dj) {
                const int i = 2 * ci - 1 + di, j = 2 * cj - 1 + dj;
                if (i > L.nx || j > L.ny) continue;
                const long c = ix(i, j, L.P);
                const float d = mg_diag(m, l, L, c);
                if (d == 0.0f) continue;
                auto tv = [&](int p, int q) -> float { return L.t[ix(p, q, L.P)]; };
                bc += L.b[c] - mg_apply(m, l, L, i, j, d, tv);
            }
        C.b[cc] = bc;
        const float dc = C.diag[cc];
        C.x[cc] = dc != 0.0f ? w * bc / dc : 0.0f;
    });
}

// coarsest level: t = cinv b (one work-group)
__global__ void __launch_bounds__(256) k_lv_dense(LvArgs a) {
    if (a.fs->done) return;
    const MGDev& mg = *a.mg;
    const MGLevelDev C = mg.lev[mg.nlev - 1];
    const int n = mg.ncoarse;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        float acc = 0.0f;
        for (int q = 0; q < n; ++q)
            acc += mg.cinv[(long)k * n + q] * C.b[ix(q / C.ny + 1, q % C.ny + 1, C.P)];
        C.t[ix(k / C.ny + 1, k % C.ny + 1, C.P)] = acc;
    }
}

// level l: y = t + wc e_parent, then x = y + w (b - A y) / d
__global__ void __launch_bounds__(256) k_lv_post(LvArgs a, int l) {
    if (a.fs->done) return;
    LV_CTX_LDS(a);
    const MGLevelDev L = a.mg->lev[l];
    const MGLevelDev C = a.mg->lev[l + 1];
    const float w = a.mg->omega, wc = a.mg->wc;
    auto corr = [&](int p, int q) -> float {
        const long c = ix(p, q, L.P);
        if (mg_diag(m, l, L, c) == 0.0f) return 0.0f;
        return L.t[c] + wc * C.t[ix((p + 1) >> 1, (q + 1) >> 1, C.P)];
    };
    lv_cell1(L.nx, L.ny, [&](int i, int j) {
        const long c = ix(i, j, L.P);
        const float d = mg_diag(m, l, L, c);
        if (d == 0.0f) return;
        const float y = corr(i, j);
        L.x[c] = y + w * (L.b[c] - mg_apply(m, l, L, i, j, d, corr)) / d;
    });
}

// Level K sums for c = 1^T b / 1^T A 1 and s = RMS(b - c A 1) (partials)
__global__ void __launch_bounds__(256) k_lv_cmode(LvArgs a, int l) {
    if (a.fs->done) return;
    LV_CTX_LDS(a);
    const MGLevelDev L = a.mg->lev[l];
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    lv_cells(L.nx, L.ny, [&](int i, int j) {
        const long c = ix(i, j, L.P);
        const float d = mg_diag(m, l, L, c);
        if (d == 0.0f) return;
        const double a1 = lv_a1(m, l, L, i, j, d);
        const double bv = L.b[c];
        s[0] += bv;
        s[1] += bv * bv;
        s[2] += bv * a1;
        s[3] += a1 * a1;
        s[4] += 1.0;
    });
    lv_partial<5>(s, a.part + LV_P_CM * LV_SLOT);
}

// one work-group: c and s of the current coarse solve; reserves a harvest
// slot when harvest_cap > 0
__global__ void __launch_bounds__(256) k_lv_cmode_fin(LvArgs a, int nb, double one_a_one,
                                                      long long harvest_cap) {
    FcgState* fs = a.fs;
    if (fs->done) return;
    double s[5];
    lv_total<5>(a.part + LV_P_CM * LV_SLOT, nb, s);
    if (threadIdx.x == 0) {
        const double cm = one_a_one > 0.0 ? s[0] / one_a_one : 0.0;
        const double bb = s[1] - 2.0 * cm * s[2] + cm * cm * s[3];
        fs->cm = cm;
        fs->s_in = s[4] > 0.0 ? sqrt(fmax(bb, 0.0) / s[4]) : 0.0;
        if (harvest_cap > 0 && fs->harvested < harvest_cap) {
            fs->hidx = fs->harvested;
            fs->harvested += 1;
        } else {
            fs->hidx = -1;
        }
    }
}

// c and s of the current coarse solve from the partials of k_lv_cmode
__device__ __forceinline__ void lv_cstats(const LvArgs& a, int nb, double one_a_one, double& cm,
                                          double& s_in) {
    double s[5];
    lv_total<5>(a.part + LV_P_CM * LV_SLOT, nb, s);
    cm = one_a_one > 0.0 ? s[0] / one_a_one : 0.0;
    const double bb = s[1] - 2.0 * cm * s[2] + cm * cm * s[3];
    s_in = s[4] > 0.0 ? sqrt(fmax(bb, 0.0) / s[4]) : 0.0;
}

// network input: the deflated, normalised level-K right-hand side on the
// padded H x W image (zero outside the fluid cells); nb = work-groups of the
// k_lv_cmode launch
__global__ void __launch_bounds__(256) k_lv_net_in(LvArgs a, int l, float* img, int H, int W,
                                                   int nb, double one_a_one) {
    FcgState* fs = a.fs;
    if (fs->skip_net) return;
    double cm, s;
    lv_cstats(a, nb, one_a_one, cm, s);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        fs->cm = cm;
        fs->s_in = s;
    }
    LV_CTX_LDS(a);
    const MGLevelDev L = a.mg->lev[l];
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long)H * W) return;
    const int i = (int)(k / W) + 1, j = (int)(k % W) + 1;
    float val = 0.0f;
    if (i <= L.nx && j <= L.ny) {
        const long c = ix(i, j, L.P);
        const float d = mg_diag(m, l, L, c);
        if (d != 0.0f && s > 0.0)
            val = (float)(((double)L.b[c] - cm * (double)lv_a1(m, l, L, i, j, d)) / s);
    }
    img[k] = val;
}

// level-K correction t = c + gain s y (gain = sigma / D; use_net = 0: t = c)
__global__ void __launch_bounds__(256) k_lv_net_out(LvArgs a, int l, const float* y, int W,
                                                    float gain, int use_net, int nb,
                                                    double one_a_one) {
    const FcgState* fs = a.fs;
    if (fs->done) return;
    double cm, s;
    lv_cstats(a, nb, one_a_one, cm, s);
    LV_CTX_LDS(a);
    const MGLevelDev L = a.mg->lev[l];
    lv_cell1(L.nx, L.ny, [&](int i, int j) {
        const long c = ix(i, j, L.P);
        if (mg_diag(m, l, L, c) == 0.0f) return;
        double t = cm;
        if (use_net && !fs->net_off) t += (double)gain * s * (double)y[(long)(i - 1) * W + (j - 1)];
        L.t[c] = (float)t;
    });
}

// Teacher solve of level K: flexible CG in FP32 vectors with FP64 dot
// products, preconditioned by V-cycles from level K down to the dense
// coarsest level, a fixed number of iterations.  Vectors on level K (pitch of
// the level): e solution, r residual, ro previous residual, d0/d1 directions,
// q = A d, bk the right-hand side.
struct TeachVec {
    float *e, *r, *ro, *d0, *d1, *q, *bk;
};

// begin: bk = b, e = 0, r = b, first V-cycle input (b, x = w b / d);
// end: t = e, b = bk
__global__ void __launch_bounds__(256) k_tch_edge(LvArgs a, int l, TeachVec tv, int end) {
    if (a.fs->done) return;
    const MGCtx m = lv_ctx(a);
    const MGLevelDev L = a.mg->lev[l];
    const float w = a.mg->omega;
    lv_cells(L.nx, L.ny, [&](int i, int j) {
        const long c = ix(i, j, L.P);
        const float d = mg_diag(m, l, L, c);
        if (d == 0.0f) return;
        if (!end) {
            const float b = L.b[c];
            tv.bk[c] = b;
            tv.e[c] = 0.0f;
            tv.r[c] = b;
            L.x[c] = w * b / d;
        } else {
            L.t[c] = tv.e[c];
            L.b[c] = tv.bk[c];
        }
    });
}

// beta = (r.z - ro.z) / rho_old (0 at it = 0), rho = r.z, z = t
__global__ void __launch_bounds__(256) k_tch_dots(LvArgs a, int l, TeachVec tv, int it) {
    FcgState* fs = a.fs;
    if (fs->done) return;
    const MGCtx m = lv_ctx(a);
    const MGLevelDev L = a.mg->lev[l];
    double s[2] = {0.0, 0.0};
    lv_cells(L.nx, L.ny, [&](int i, int j) {
        const long c = ix(i, j, L.P);
        if (mg_diag(m, l, L, c) == 0.0f) return;
        const double z = L.t[c];
        s[0] += (double)tv.r[c] * z;
        if (it > 0) s[1] += (double)tv.ro[c] * z;
    });
    if (lv_reduce<2>(s, a.part, &fs->ticket[1]) && threadIdx.x == 0) {
        fs->ibeta = it == 0 ? 0.0 : (s[0] - s[1]) / fs->irho_old;
        fs->irho = s[0];
    }
}

// d = z + beta d_old, q = A d, alpha = rho / d.q
__global__ void __launch_bounds__(256) k_tch_dir(LvArgs a, int l, TeachVec tv, int it) {
    FcgState* fs = a.fs;
    if (fs->done) return;
    const MGCtx m = lv_ctx(a);
    const MGLevelDev L = a.mg->lev[l];
    const float beta = (float)fs->ibeta;
    float* dn = (it & 1) ? tv.d1 : tv.d0;
    const float* dd = (it & 1) ? tv.d0 : tv.d1;
    auto dirv = [&](int p, int q) -> float {
        const long c = ix(p, q, L.P);
        return it == 0 ? L.t[c] : L.t[c] + beta * dd[c];
    };
    double pq[1] = {0.0};
    lv_cells(L.nx, L.ny, [&](int i, int j) {
        const long c = ix(i, j, L.P);
        const float d = mg_diag(m, l, L, c);
        if (d == 0.0f) return;
        const float dc = dirv(i, j);
        const float q = mg_apply(m, l, L, i, j, d, dirv);
        dn[c] = dc;
        tv.q[c] = q;
        pq[0] += (double)dc * (double)q;
    });
    if (lv_reduce<1>(pq, a.part, &fs->ticket[2]) && threadIdx.x == 0)
        fs->ialpha = pq[0] > 0.0 ? fs->irho / pq[0] : 0.0;
}

// e += alpha d, ro = r, r -= alpha q; next V-cycle input b = r, x = w r / d
__global__ void __launch_bounds__(256) k_tch_update(LvArgs a, int l, TeachVec tv, int it) {
    FcgState* fs = a.fs;
    if (fs->done) return;
    const MGCtx m = lv_ctx(a);
    const MGLevelDev L = a.mg->lev[l];
    const float w = a.mg->omega;
    const float alpha = (float)fs->ialpha;
    const float* dn = (it & 1) ? tv.d1 : tv.d0;
    lv_cells(L.nx, L.ny, [&](int i, int j) {
        const long c = ix(i, j, L.P);
        const float d = mg_diag(m, l, L, c);
        if (d == 0.0f) return;
        tv.e[c] += alpha * dn[c];
        const float rv = tv.r[c];
        tv.ro[c] = rv;
        const float rn = rv - alpha * tv.q[c];
        tv.r[c] = rn;
        L.b[c] = rn;
        L.x[c] = w * rn / d;
    });
    if (blockIdx.x == 0 && threadIdx.x == 0) fs->irho_old = fs->irho;
}

// store the pair (network input, (t - c) D / s) in harvest slot fs->hidx
__global__ void __launch_bounds__(256) k_lv_harvest(LvArgs a, int l, float* hin, float* hout,
                                                    int H, int W, float D) {
    const FcgState* fs = a.fs;
    if (fs->done || fs->hidx < 0) return;
    const MGCtx m = lv_ctx(a);
    const MGLevelDev L = a.mg->lev[l];
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long)H * W) return;
    const int i = (int)(k / W) + 1, j = (int)(k % W) + 1;
    float vin = 0.0f, vout = 0.0f;
    const double s = fs->s_in;
    if (i <= L.nx && j <= L.ny && s > 0.0) {
        const long c = ix(i, j, L.P);
        const float d = mg_diag(m, l, L, c);
        if (d != 0.0f) {
            vin = (float)(((double)L.b[c] - fs->cm * (double)lv_a1(m, l, L, i, j, d)) / s);
            vout = (float)(((double)L.t[c] - fs->cm) * D / s);
        }
    }
    const long base = fs->hidx * (long)H * W;
    hin[base + k] = vin;
    hout[base + k] = vout;
}

// ---- the legs of the V-cycle above the coarse end, one launch per level and leg
// A work-group takes tiles of TI x TJ cells of level l (TI, TJ even, so a
// tile holds whole 2 x 2 aggregates) and stages them in LDS with a ring of NU
// cells (rows wrapped when periodic; cells outside the arrays count as
// solid), so that the NU damped-Jacobi sweeps of a leg run back to back with
// barriers instead of launches, each sweep on a ring one cell narrower:
//   down-leg (k_lv_down): x = w b / d (the first sweep from zero, as the
//     producers of b store it), NU - 1 sweeps t = x + w (b - A x) / d, and
//     the residual b - A t summed over each aggregate into b of level l + 1
//     (x = w b / d there);
//   up-leg (k_lv_up): y = t + wc t_(l+1) (the prolongation of level l + 1's
//     correction onto the down-leg's t), then NU sweeps; on level 0 with the
//     partial dot products r.z and r_old.z of the flexible CG.
// With NU = 2 the values are those of k_lv_smooth / k_lv_restrict / k_lv_post
// / k_lv_smooth_dots (the V-cycle of the persistent multigrid solver),
// expression for expression: a ring cell is recomputed from the same inputs by
// the same FP32 arithmetic.  NU > 2 is a stronger preconditioner, still
// symmetric (as many sweeps down as up).  The down-leg's t goes to `tdown`
// (the level's x array: no later kernel reads x there, k_lv_coarse and the
// next down-leg start from b), where the up-leg reads it back: the up-leg
// writes t, which the work-groups of neighbouring tiles read in their rings,
// so it must not be the array it reads.
struct LegArgs {
    MGLevelDev L, C;          // level l and level l + 1
    float* tdown;             // down-leg: t of level l goes here
    const float* tup;         // up-leg: the down-leg's t of level l
    const uint8_t* code;      // level 0: cell codes (pitch L.P)
    const float* diag0;       // level 0: diagonal table (64 floats)
    const int* done;          // the solve has finished
    const int* first;         // the flexible CG (re)starts: no r_old.z
    const double *r, *rold;   // level-0 up-leg: FP64 residuals (pitch L.P)
    double* part;             // level-0 up-leg: dot partials, 2 per work-group
    float w, wc, ax, ay;
    int periodic;
};

namespace {

// row of a ring cell: wrapped into 1..nx when periodic (mg_apply's iE / iW);
// otherwise unchanged (0 and nx + 1 are the ghost rows)
__device__ __forceinline__ int leg_row(int i, int nx, int periodic) {
    if (!periodic) return i;
    int r = (i - 1) % nx;
    if (r < 0) r += nx;
    return r + 1;
}

// (A v) at LDS cell k of a tile of row width W: mg_apply's expressions
template <bool LEV0>
__device__ __forceinline__ float leg_apply(const LegArgs& a, int k, int W, float d, const float* v,
                                           const float* e, const float* n) {
    if (LEV0) return d * v[k] - a.ax * (v[k + W] + v[k - W]) - a.ay * (v[k + 1] + v[k - 1]);
    return d * v[k] - e[k] * v[k + W] - e[k - W] * v[k - W] - n[k] * v[k + 1] - n[k - 1] * v[k - 1];
}

// this work-group's tiles [t0, t1) of level L: contiguous runs of tiles
// (row-major over the level) per XCD
template <int TI, int TJ>
__device__ __forceinline__ void leg_tiles(const MGLevelDev& L, int& ntj, int& t0, int& t1) {
    ntj = (L.ny + TJ - 1) / TJ;
    const long nt = (long)((L.nx + TI - 1) / TI) * ntj;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    t0 = (int)((nt * b) / gridDim.x);
    t1 = (int)((nt * (b + 1)) / gridDim.x);
}

// The staging of a tile's LDS region (the tile and its ring of R cells), 256
// threads: every global load of the thread's cells is issued before any is
// used.  Per cell: b, the diagonal (level 0: from the code), the couplings
// (levels >= 1) and either x = w b / d (down-leg) or the prolongated
// correction y = t + wc t_(l+1) (up-leg; k_lv_post's expression, 0 off the
// fluid).  (ib, jb): the level cell of LDS cell (0, 0).  With `gate` (the
// first tile of a coarse level's launch) the solve's finished flag is tested
// only once the tile's loads are in flight (false: finished, nothing is
// staged).
template <bool LEV0, bool UP, int TI, int TJ, int R, int BS = 256>
__device__ __forceinline__ bool leg_stage(const LegArgs& a, int ib, int jb, const float* s_d0,
                                          float* sb, float* sd, float* se, float* sn, float* sv,
                                          const int* gate) {
    constexpr int W = TJ + 2 * R, N = (TI + 2 * R) * W, LA = (N + BS - 1) / BS;
    const MGLevelDev& L = a.L;
    const MGLevelDev& C = a.C;
    float rb[LA], rd[LA], re[LA], rn[LA], rt[LA], rc[LA];
    int rcode[LA];
    bool ok[LA];
#pragma unroll
    for (int u = 0; u < LA; ++u) {
        const int e = (int)threadIdx.x + u * BS;
        const int r = e / W, q = e - r * W;
        const int i = leg_row(ib + r, L.nx, a.periodic), j = jb + q;
        ok[u] = e < N && i >= 0 && i <= L.nx + 1 && j >= 0 && j <= L.ny + 1;
        const long c = ok[u] ? ix(i, j, L.P) : 0;
        rb[u] = ok[u] ? L.b[c] : 0.0f;
        if (LEV0) {
            rcode[u] = ok[u] ? a.code[c] : 0;
        } else {
            rd[u] = ok[u] ? L.diag[c] : 0.0f;
            re[u] = ok[u] ? L.cE[c] : 0.0f;
            rn[u] = ok[u] ? L.cN[c] : 0.0f;
        }
        if (UP) {
            rt[u] = ok[u] ? a.tup[c] : 0.0f;
            rc[u] = ok[u] ? C.t[ix((i + 1) >> 1, (j + 1) >> 1, C.P)] : 0.0f;
        }
    }
    if (gate && *gate) return false;
#pragma unroll
    for (int u = 0; u < LA; ++u) {
        const int e = (int)threadIdx.x + u * BS;
        if (e >= N) continue;
        const float d = LEV0 ? s_d0[rcode[u] & 63] : rd[u];
        sb[e] = rb[u];
        sd[e] = d;
        if (!LEV0) {
            se[e] = re[u];
            sn[e] = rn[u];
        }
        if (UP) {
            const float v = rt[u] + a.wc * rc[u];
            sv[e] = d == 0.0f ? 0.0f : v;
        } else {
            sv[e] = d != 0.0f ? a.w * rb[u] / d : 0.0f;
        }
    }
    return true;
}

// dst = src + w (b - A src) / d on the tile and its ring of RHO cells (0 off
// the fluid), in an LDS region with a ring of R cells
// (tid: the thread's index, see leg_tid)
template <bool LEV0, int TI, int TJ, int R, int RHO, int BS = 256>
__device__ __forceinline__ void leg_sweep(const LegArgs& a, const float* sb, const float* sd,
                                          const float* se, const float* sn, const float* src,
                                          float* dst, int tid) {
    constexpr int W = TJ + 2 * R, W1 = TJ + 2 * RHO, N1 = (TI + 2 * RHO) * W1;
    constexpr int LS = (N1 + BS - 1) / BS, O = R - RHO;
#pragma unroll
    for (int u = 0; u < LS; ++u) {
        const int e = tid + u * BS;
        if (e >= N1) continue;
        const int r = e / W1, q = e - r * W1;
        const int k = (r + O) * W + q + O;
        const float d = sd[k];
        dst[k] = d != 0.0f ? src[k] + a.w * (sb[k] - leg_apply<LEV0>(a, k, W, d, src, se, sn)) / d
                           : 0.0f;
    }
}

// NS sweeps on the rings RHO, RHO - 1, .. alternating between p and q, each
// followed by a barrier; the buffer that holds the last result
template <bool LEV0, int TI, int TJ, int R, int RHO, int NS, int BS = 256>
__device__ __forceinline__ float* leg_sweeps(const LegArgs& a, const float* sb, const float* sd,
                                             const float* se, const float* sn, float* p, float* q,
                                             int tid = (int)threadIdx.x) {
    if constexpr (NS == 0) {
        return p;
    } else {
        leg_sweep<LEV0, TI, TJ, R, RHO, BS>(a, sb, sd, se, sn, p, q, tid);
        __syncthreads();
        return leg_sweeps<LEV0, TI, TJ, R, RHO - 1, NS - 1, BS>(a, sb, sd, se, sn, q, p, tid);
    }
}

// The thread's index through an empty asm statement, taken inside a tile
// loop: the compiler can then not hoist the sweeps' LDS offsets out of the
// loop, where they stayed live through it and were spilled (the up-legs:
// 146 -> 82 registers, 15.3 -> 13.2 us on level 0); the down-legs are
// faster with them hoisted
__device__ __forceinline__ int leg_tid() {
    int tid = (int)threadIdx.x;
    asm volatile("" : "+v"(tid));
    return tid;
}

}  // namespace

template <bool LEV0, int NU, int TI, int TJ>
__global__ void __launch_bounds__(256) k_lv_down(LegArgs a) {
    constexpr int R = NU, W = TJ + 2 * R, N = (TI + 2 * R) * W;
    __shared__ float sb[N], sd[N], sx[N], st[N];
    __shared__ float se[LEV0 ? 1 : N], sn[LEV0 ? 1 : N];
    __shared__ float s_d0[64];
    // level 0: the diagonal table's load in flight with the flag's; the coarse
    // levels test the flag once the first tile's loads are in flight
    // (measured: 0.25 us less per launch there, 0.5 us more on level 0)
    if (LEV0 && threadIdx.x < 64) s_d0[threadIdx.x] = a.diag0[threadIdx.x];
    if (LEV0) {
        if (*a.done) return;
        __syncthreads();
    }
    const MGLevelDev& L = a.L;
    const MGLevelDev& C = a.C;
    const float w = a.w;
    int ntj, t0, t1;
    leg_tiles<TI, TJ>(L, ntj, t0, t1);
    for (int tile = t0; tile < t1; ++tile) {
        const int ib = (tile / ntj) * TI + 1 - R;   // level row of LDS row 0
        const int jb = (tile % ntj) * TJ + 1 - R;
        if (!leg_stage<LEV0, false, TI, TJ, R>(a, ib, jb, s_d0, sb, sd, se, sn, sx,
                                                !LEV0 && tile == t0 ? a.done : nullptr))   // x = w b / d
            return;
        __syncthreads();
        // t on the tile and its first ring
        const float* tt = leg_sweeps<LEV0, TI, TJ, R, R - 1, NU - 1>(a, sb, sd, se, sn, sx, st);
        // the residual restricted: one aggregate of level l + 1 per thread
        const int ci0 = (tile / ntj) * (TI / 2) + 1, cj0 = (tile % ntj) * (TJ / 2) + 1;
        constexpr int NC = (TI / 2) * (TJ / 2), LC = (NC + 255) / 256;
#pragma unroll
        for (int u = 0; u < LC; ++u) {
            const int e = (int)threadIdx.x + u * 256;
            const int cr = e / (TJ / 2), cs = e - cr * (TJ / 2);
            const int ci = ci0 + cr, cj = cj0 + cs;
            if (e >= NC || ci > C.nx || cj > C.ny) continue;
            float bc = 0.0f;
            for (int di = 0; di < 2; ++di)
                for (int dj = 0; dj < 2; ++dj) {
                    const int i = 2 * ci - 1 + di, j = 2 * cj - 1 + dj;
                    if (i > L.nx || j > L.ny) continue;
                    const int k = (R + 2 * cr + di) * W + R + 2 * cs + dj;
                    const float d = sd[k];
                    if (d == 0.0f) continue;
                    bc += sb[k] - leg_apply<LEV0>(a, k, W, d, tt, se, sn);
                }
            const long cc = ix(ci, cj, C.P);
            C.b[cc] = bc;
            const float dc = C.diag[cc];
            C.x[cc] = dc != 0.0f ? w * bc / dc : 0.0f;
        }
        constexpr int NI = TI * TJ, LI = (NI + 255) / 256;
#pragma unroll
        for (int u = 0; u < LI; ++u) {
            const int e = (int)threadIdx.x + u * 256;
            const int r = e / TJ, q = e - r * TJ;
            const int i = ib + R + r, j = jb + R + q;
            const int k = (r + R) * W + q + R;
            if (e < NI && i <= L.nx && j <= L.ny && sd[k] != 0.0f) a.tdown[ix(i, j, L.P)] = tt[k];
        }
        __syncthreads();   // the next tile overwrites the LDS arrays
    }
}

// the fine level's 16 x 64 tiles with 3 sweeps held to 128 registers: 4
// work-groups per CU make the 1024 tiles of a 1024^2 level one round
// the fine level's 16 x 64 tiles with 3 sweeps held to 128 registers (a few
// spilled in the staging): 4 work-groups per CU make the 1024 tiles of a
// 1024^2 level one round (18.3 -> 15.1 us per launch)
template <bool LEV0, bool DOTS, int NU, int TI, int TJ>
__global__ void __launch_bounds__(256, (LEV0 && TI == 16 && NU == 3) ? 4 : 1) k_lv_up(LegArgs a) {
    constexpr int R = NU, W = TJ + 2 * R, N = (TI + 2 * R) * W;
    __shared__ float sb[N], sd[N], sy[N], sx[N];
    __shared__ float se[LEV0 ? 1 : N], sn[LEV0 ? 1 : N];
    __shared__ float s_d0[64];
    // level 0: the diagonal table's load in flight with the flag's; the coarse
    // levels test the flag once the first tile's loads are in flight
    if (LEV0 && threadIdx.x < 64) s_d0[threadIdx.x] = a.diag0[threadIdx.x];
    if (LEV0) {
        if (*a.done) return;
        __syncthreads();
    }
    const MGLevelDev& L = a.L;
    const float w = a.w;
    const bool first = DOTS ? *a.first != 0 : true;
    double s2[2] = {0.0, 0.0};
    int ntj, t0, t1;
    leg_tiles<TI, TJ>(L, ntj, t0, t1);
    for (int tile = t0; tile < t1; ++tile) {
        const int ib = (tile / ntj) * TI + 1 - R;
        const int jb = (tile % ntj) * TJ + 1 - R;
        if (!leg_stage<LEV0, true, TI, TJ, R>(a, ib, jb, s_d0, sb, sd, se, sn, sy,
                                               !LEV0 && tile == t0 ? a.done : nullptr))   // y
            return;
        constexpr int NI = TI * TJ, LI = (NI + 255) / 256;
        double rr[LI], ro[LI];
#pragma unroll
        for (int u = 0; u < LI; ++u) {   // the residuals' loads, in flight during the sweeps
            const int e = (int)threadIdx.x + u * 256;
            const int r = e / TJ, q = e - r * TJ;
            const int i = ib + R + r, j = jb + R + q;
            const bool in = e < NI && i <= L.nx && j <= L.ny;
            const long c = in ? ix(i, j, L.P) : 0;
            rr[u] = DOTS && in ? a.r[c] : 0.0;
            ro[u] = DOTS && in && !first ? a.rold[c] : 0.0;
        }
        __syncthreads();
        // NU - 1 sweeps down to the first ring, the last one on the tile below
        const float* xx = leg_sweeps<LEV0, TI, TJ, R, R - 1, NU - 1>(a, sb, sd, se, sn, sy, sx, leg_tid());
#pragma unroll
        for (int u = 0; u < LI; ++u) {
            const int e = (int)threadIdx.x + u * 256;
            const int r = e / TJ, q = e - r * TJ;
            const int i = ib + R + r, j = jb + R + q;
            const int k = (r + R) * W + q + R;
            if (e >= NI || i > L.nx || j > L.ny) continue;
            const float d = sd[k];
            if (d == 0.0f) continue;
            const float z = xx[k] + w * (sb[k] - leg_apply<LEV0>(a, k, W, d, xx, se, sn)) / d;
            L.t[ix(i, j, L.P)] = z;
            if (DOTS) {
                s2[0] += rr[u] * (double)z;
                if (!first) s2[1] += ro[u] * (double)z;
            }
        }
        __syncthreads();
    }
    if (DOTS) lv_partial<2>(s2, a.part + LV_P_DOTS * LV_SLOT);
}

// ---- two coarse levels per leg launch -------------------------------------
// Levels l (F, not level 0) and l + 1 (M) whose legs k_lv_down / k_lv_up
// would take one launch each, in one launch per leg: one TI x TJ tile of M per
// work-group (LEG2_DOWN_BS / LEG2_UP_BS threads).
// Every cell value a work-group needs from the other level is recomputed in
// its LDS from the same inputs by the same FP32 expressions as the
// single-level kernels (the fields are bit for bit those of the two
// launches): the down-leg smooths F on the region under the M tile and its
// ring of R cells (plus F's own ring), restricts F's residual into M's b
// there, then runs M's leg; the up-leg runs M's leg on the tile and a ring of
// (R + 1) / 2 cells, whose t the F cells of the tile under it and their ring
// prolongate.  The second level's global loads are issued with the first's.
// Closed (not periodic) domains with nx(F) = 2 nx(M), ny(F) = 2 ny(M) only
// (the host checks).
constexpr int LEG2_DOWN_BS = 1024, LEG2_UP_BS = 1024;

template <int NU, int TI, int TJ>
__global__ void __launch_bounds__(LEG2_DOWN_BS) k_lv_down2(LegArgs f, LegArgs m) {
    constexpr int R = NU, BS = LEG2_DOWN_BS;
    // F: the region under the M tile and its ring R, plus F's ring R
    constexpr int FI = 2 * (TI + 2 * R), FJ = 2 * (TJ + 2 * R);
    constexpr int FW = FJ + 2 * R;
    constexpr int MW = TJ + 2 * R, MN = (TI + 2 * R) * MW, LM = (MN + BS - 1) / BS;
    __shared__ float fb[(FI + 2 * R) * FW], fd[(FI + 2 * R) * FW], fe[(FI + 2 * R) * FW],
        fn[(FI + 2 * R) * FW], fx[(FI + 2 * R) * FW], ft[(FI + 2 * R) * FW];
    __shared__ float mb[MN], md[MN], me[MN], mn[MN], mx[MN], mt[MN];
    const MGLevelDev& F = f.L;
    const MGLevelDev& M = m.L;
    const MGLevelDev& C = m.C;
    const float w = f.w;
    int ntj, t0, t1;
    leg_tiles<TI, TJ>(M, ntj, t0, t1);
    for (int tile = t0; tile < t1; ++tile) {
        const int ci0 = (tile / ntj) * TI + 1, cj0 = (tile % ntj) * TJ + 1;   // M tile origin
        const int mib = ci0 - R, mjb = cj0 - R;                             // M region origin
        // M's coefficients on its region (leg_stage's loads), in flight with F's stage
        float pd[LM], pe[LM], pn[LM];
#pragma unroll
        for (int u = 0; u < LM; ++u) {
            const int e = (int)threadIdx.x + u * BS;
            const int r = e / MW, q = e - r * MW;
            const int i = mib + r, j = mjb + q;
            const bool ok = e < MN && i >= 0 && i <= M.nx + 1 && j >= 0 && j <= M.ny + 1;
            const long c = ok ? ix(i, j, M.P) : 0;
            pd[u] = ok ? M.diag[c] : 0.0f;
            pe[u] = ok ? M.cE[c] : 0.0f;
            pn[u] = ok ? M.cN[c] : 0.0f;
        }
        // F cell of F-LDS cell (0, 0): the first child row of M row mib, less F's ring
        const int fib = 2 * mib - 1 - R, fjb = 2 * mjb - 1 - R;
        if (!leg_stage<false, false, FI, FJ, R, BS>(f, fib, fjb, nullptr, fb, fd, fe, fn, fx,
                                                     tile == t0 ? f.done : nullptr))
            return;
        __syncthreads();
        const float* tf =
            leg_sweeps<false, FI, FJ, R, R - 1, NU - 1, BS>(f, fb, fd, fe, fn, fx, ft, leg_tid());
        // F's residual restricted: M's b on the M region (k_lv_down's loop),
        // M's first iterate x = w b / d (leg_stage's expression)
#pragma unroll
        for (int u = 0; u < LM; ++u) {
            const int e = (int)threadIdx.x + u * BS;
            if (e >= MN) continue;
            const int cr = e / MW, cs = e - cr * MW;
            const int ci = mib + cr, cj = mjb + cs;
            float bc = 0.0f;
            if (ci >= 1 && cj >= 1 && ci <= M.nx && cj <= M.ny) {
                for (int di = 0; di < 2; ++di)
                    for (int dj = 0; dj < 2; ++dj) {
                        const int i = 2 * ci - 1 + di, j = 2 * cj - 1 + dj;
                        if (i > F.nx || j > F.ny) continue;
                        const int k = (R + 2 * cr + di) * FW + R + 2 * cs + dj;
                        const float d = fd[k];
                        if (d == 0.0f) continue;
                        bc += fb[k] - leg_apply<false>(f, k, FW, d, tf, fe, fn);
                    }
            }
            const float d = pd[u];
            mb[e] = bc;
            md[e] = d;
            me[e] = pe[u];
            mn[e] = pn[u];
            mx[e] = d != 0.0f ? m.w * bc / d : 0.0f;
        }
        __syncthreads();
        const float* tm =
            leg_sweeps<false, TI, TJ, R, R - 1, NU - 1, BS>(m, mb, md, me, mn, mx, mt, leg_tid());
        // M's residual restricted into level l + 2 (k_lv_down's loop)
        const int ki0 = (tile / ntj) * (TI / 2) + 1, kj0 = (tile % ntj) * (TJ / 2) + 1;
        constexpr int NC = (TI / 2) * (TJ / 2), LC = (NC + BS - 1) / BS;
#pragma unroll
        for (int u = 0; u < LC; ++u) {
            const int e = (int)threadIdx.x + u * BS;
            const int cr = e / (TJ / 2), cs = e - cr * (TJ / 2);
            const int ci = ki0 + cr, cj = kj0 + cs;
            if (e >= NC || ci > C.nx || cj > C.ny) continue;
            float bc = 0.0f;
            for (int di = 0; di < 2; ++di)
                for (int dj = 0; dj < 2; ++dj) {
                    const int i = 2 * ci - 1 + di, j = 2 * cj - 1 + dj;
                    if (i > M.nx || j > M.ny) continue;
                    const int k = (R + 2 * cr + di) * MW + R + 2 * cs + dj;
                    const float d = md[k];
                    if (d == 0.0f) continue;
                    bc += mb[k] - leg_apply<false>(m, k, MW, d, tm, me, mn);
                }
            const long cc = ix(ci, cj, C.P);
            C.b[cc] = bc;
            const float dc = C.diag[cc];
            C.x[cc] = dc != 0.0f ? w * bc / dc : 0.0f;
        }
        // M's b (its up-leg stages it) and both levels' t under the tile
#pragma unroll
        for (int u = 0; u < (TI * TJ + BS - 1) / BS; ++u) {
            const int e = (int)threadIdx.x + u * BS;
            const int r = e / TJ, q = e - r * TJ;
            const int i = ci0 + r, j = cj0 + q;
            const int k = (r + R) * MW + q + R;
            if (e >= TI * TJ || i > M.nx || j > M.ny) continue;
            M.b[ix(i, j, M.P)] = mb[k];
            if (md[k] != 0.0f) m.tdown[ix(i, j, M.P)] = tm[k];
        }
#pragma unroll
        for (int u = 0; u < (4 * TI * TJ + BS - 1) / BS; ++u) {
            const int e = (int)threadIdx.x + u * BS;
            const int r = e / (2 * TJ), q = e - r * (2 * TJ);
            const int i = 2 * ci0 - 1 + r, j = 2 * cj0 - 1 + q;
            const int k = (R + 2 * R + r) * FW + R + 2 * R + q;
            if (e >= 4 * TI * TJ || i > F.nx || j > F.ny) continue;
            if (fd[k] != 0.0f) f.tdown[ix(i, j, F.P)] = tf[k];
        }
        __syncthreads();   // the next tile overwrites the LDS arrays
    }
}

template <int NU, int TI, int TJ>
__global__ void __launch_bounds__(LEG2_UP_BS) k_lv_up2(LegArgs m, LegArgs f) {
    constexpr int R = NU, RP = (R + 1) / 2, BS = LEG2_UP_BS;
    // M: the tile and a ring of RP cells as the "tile", with M's own ring R
    constexpr int EI = TI + 2 * RP, EJ = TJ + 2 * RP;
    constexpr int MW = EJ + 2 * R, MN = (EI + 2 * R) * MW;
    // F: the 2 TI x 2 TJ cells under the M tile, with F's ring R
    constexpr int FI = 2 * TI, FJ = 2 * TJ, FW = FJ + 2 * R, FN = (FI + 2 * R) * FW;
    constexpr int LA = (FN + BS - 1) / BS;
    __shared__ float mb[MN], md[MN], me[MN], mn[MN], my[MN], mx[MN];
    __shared__ float fb[FN], fd[FN], fe[FN], fn[FN], fy[FN], fx[FN];
    const MGLevelDev& F = f.L;
    const MGLevelDev& M = m.L;
    int ntj, t0, t1;
    leg_tiles<TI, TJ>(M, ntj, t0, t1);
    for (int tile = t0; tile < t1; ++tile) {
        const int ci0 = (tile / ntj) * TI + 1, cj0 = (tile % ntj) * TJ + 1;
        const int mib = ci0 - RP - R, mjb = cj0 - RP - R;
        const int fib = 2 * ci0 - 1 - R, fjb = 2 * cj0 - 1 - R;
        // F's stage loads (k_lv_up's), in flight with M's
        float rb[LA], rd[LA], re[LA], rn[LA], rt[LA];
        int kc[LA];
        bool ok[LA];
#pragma unroll
        for (int u = 0; u < LA; ++u) {
            const int e = (int)threadIdx.x + u * BS;
            const int r = e / FW, q = e - r * FW;
            const int i = fib + r, j = fjb + q;
            ok[u] = e < FN && i >= 0 && i <= F.nx + 1 && j >= 0 && j <= F.ny + 1;
            const long c = ok[u] ? ix(i, j, F.P) : 0;
            rb[u] = ok[u] ? F.b[c] : 0.0f;
            rd[u] = ok[u] ? F.diag[c] : 0.0f;
            re[u] = ok[u] ? F.cE[c] : 0.0f;
            rn[u] = ok[u] ? F.cN[c] : 0.0f;
            rt[u] = ok[u] ? f.tup[c] : 0.0f;
            // the parent's cell in M's LDS region (inside it for every F cell staged)
            kc[u] = (((i + 1) >> 1) - mib) * MW + (((j + 1) >> 1) - mjb);
        }
        // M: its up-leg on the extended tile (k_lv_up's stage, sweeps, last sweep)
        if (!leg_stage<false, true, EI, EJ, R, BS>(m, mib, mjb, nullptr, mb, md, me, mn, my,
                                                    tile == t0 ? m.done : nullptr))
            return;
        __syncthreads();
        const float* xm =
            leg_sweeps<false, EI, EJ, R, R - 1, NU - 1, BS>(m, mb, md, me, mn, my, mx, leg_tid());
        float* tm = xm == my ? mx : my;   // the last sweep's result goes to the other buffer
        constexpr int NE = EI * EJ;
#pragma unroll
        for (int u = 0; u < (NE + BS - 1) / BS; ++u) {
            const int e = (int)threadIdx.x + u * BS;
            if (e >= NE) continue;
            const int r = e / EJ, q = e - r * EJ;
            const int k = (r + R) * MW + q + R;
            const float d = md[k];
            tm[k] = d != 0.0f ? xm[k] + m.w * (mb[k] - leg_apply<false>(m, k, MW, d, xm, me, mn)) / d
                              : 0.0f;
        }
        __syncthreads();
        // F: k_lv_up's stage with the prolongated M t from LDS
#pragma unroll
        for (int u = 0; u < LA; ++u) {
            const int e = (int)threadIdx.x + u * BS;
            if (e >= FN) continue;
            const float d = rd[u];
            fb[e] = rb[u];
            fd[e] = d;
            fe[e] = re[u];
            fn[e] = rn[u];
            const float rc = ok[u] ? tm[kc[u]] : 0.0f;
            const float v = rt[u] + f.wc * rc;
            fy[e] = d == 0.0f ? 0.0f : v;
        }
        __syncthreads();
        const float* xf =
            leg_sweeps<false, FI, FJ, R, R - 1, NU - 1, BS>(f, fb, fd, fe, fn, fy, fx, leg_tid());
#pragma unroll
        for (int u = 0; u < (FI * FJ + BS - 1) / BS; ++u) {
            const int e = (int)threadIdx.x + u * BS;
            const int r = e / FJ, q = e - r * FJ;
            const int i = fib + R + r, j = fjb + R + q;
            const int k = (r + R) * FW + q + R;
            if (e >= FI * FJ || i > F.nx || j > F.ny) continue;
            const float d = fd[k];
            if (d == 0.0f) continue;
            F.t[ix(i, j, F.P)] = xf[k] + f.w * (fb[k] - leg_apply<false>(f, k, FW, d, xf, fe, fn)) / d;
        }
        __syncthreads();
    }
}

// ---- flexible CG steps on the fine level (z = lev[0].t) ----
// beta = (r.z - r_old.z) / rho_old (0 at a (re)start) from the dot partials of
// nb work-groups; d = z + beta d_old, q = A d (neighbours' d rebuilt from z and
// d_old), partial d.q
__global__ void __launch_bounds__(512) k_fcg_dir(LvArgs a, FcgVec v, int nb) {
    FcgState* fs = a.fs;
    // the flag, the table and the partials in flight together (the partial
    // sums' barriers also publish the table)
    const int done = fs->done;
    __shared__ double s_lut[64];
    if (threadIdx.x < 64) s_lut[threadIdx.x] = v.lut[threadIdx.x];
    double dots[2];
    lv_total<2>(a.part + LV_P_DOTS * LV_SLOT, nb, dots);
    if (done) return;
    const bool first = fs->first;
    const double beta = first ? 0.0 : (dots[0] - dots[1]) / fs->rho_old;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        fs->rho = dots[0];
        fs->beta = beta;
    }
    const float* z = a.mg->lev[0].t;
    const int ny = a.mg->lev[0].ny;
    const int k = fs->k;
    double* dnew = (k & 1) ? v.d1 : v.d0;
    const double* dold = (k & 1) ? v.d0 : v.d1;
    double pq[1] = {0.0};
    struct D {
        int li;
        double dc, q;
    };
    // two variants (restart or not) keep the loads free of branches
    auto sweep = [&](auto first_c) {
        constexpr bool FIRST = decltype(first_c)::value;
        lv_cells_batch<4>(
            v.nx, ny,
            [&](int i, int j) {
                const long c = ix(i, j, v.pc);
                D o;
                o.li = a.code[c] & 63;
                const int iE = (a.periodic && i == v.nx) ? 1 : i + 1;
                const int iW = (a.periodic && i == 1) ? v.nx : i - 1;
                auto dirv = [&](long cc) -> double {
                    const double zc = z[cc];
                    return FIRST ? zc : zc + beta * dold[cc];
                };
                o.dc = dirv(c);
                o.q = s_lut[o.li] * o.dc - v.ax * (dirv(ix(iE, j, v.pc)) + dirv(ix(iW, j, v.pc))) -
                      v.ay * (dirv(ix(i, j + 1, v.pc)) + dirv(ix(i, j - 1, v.pc)));
                return o;
            },
            [&](int i, int j, const D& o) {
                if (!o.li) return;
                const long c = ix(i, j, v.pc);
                dnew[c] = o.dc;
                v.q[c] = o.q;
                pq[0] += o.dc * o.q;
            });
    };
    if (first) sweep(std::true_type{});
    else sweep(std::false_type{});
    lv_partial<1>(pq, a.part + LV_P_PQ * LV_SLOT);
}

// alpha = rho / d.q; x += alpha d, r_old = r, r -= alpha q; next V-cycle
// input b0 = r, x0 = w r / d; partial r.r
__global__ void __launch_bounds__(512) k_fcg_update(LvArgs a, FcgVec v, int nb) {
    FcgState* fs = a.fs;
    // the flag, the diagonal table and the partials in flight together (the
    // partial sums' barriers also publish the table)
    const int done = fs->done;
    __shared__ float s_d0_[64];
    if (threadIdx.x < 64) s_d0_[threadIdx.x] = a.diag0[threadIdx.x];
    double pq[1];
    lv_total<1>(a.part + LV_P_PQ * LV_SLOT, nb, pq);
    if (done) return;
    const double alpha = fs->rho / pq[0];
    if (!(pq[0] > 0.0) || !(fabs(alpha) < 1e300)) {   // breakdown or non-finite
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            fs->done = 1;
            fs->info = -1;
        }
        return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) fs->alpha = alpha;
    const MGCtx m{a.mg, a.code, s_d0_, a.periodic, a.ax, a.ay};
    const MGLevelDev L0 = a.mg->lev[0];
    const float w = a.mg->omega;
    const int k = fs->k;
    const double* dnew = (k & 1) ? v.d1 : v.d0;
    double s[1] = {0.0};
    struct Up {
        int li;
        double x, r, rn;
    };
    lv_cells_batch<4>(
        v.nx, L0.ny,
        [&](int i, int j) {
            const long c = ix(i, j, v.pc);
            Up o;
            o.li = a.code[c] & 63;
            o.x = v.x[c] + alpha * dnew[c];
            o.r = v.r[c];
            o.rn = o.r - alpha * v.q[c];
            return o;
        },
        [&](int i, int j, const Up& o) {
            if (!o.li) return;
            const long c = ix(i, j, v.pc);
            v.x[c] = o.x;
            v.rold[c] = o.r;
            v.r[c] = o.rn;
            s[0] += o.rn * o.rn;
            const float rf = (float)o.rn;
            L0.b[c] = rf;
            if (a.store_x0) L0.x[c] = w * rf / m.diag0[o.li];
        });
    lv_partial<1>(s, a.part + LV_P_RR * LV_SLOT);
}

// End of an iteration: every work-group adds the r.r partials of
// k_fcg_update and applies the recursive-residual test (work-group 0 records
// it); once it passes (or at maxiter) the true residual r = b - A x is formed
// with the next V-cycle input, and its partial r.r stored for k_fcg_true_fin
// in the d.q slot (free after k_fcg_update): the r.r slot is still being read
// by work-groups that start later.
__global__ void __launch_bounds__(512) k_fcg_true(LvArgs a, FcgVec v, int nb) {
    FcgState* fs = a.fs;
    const int done = fs->done;   // in flight with the partials
    double rr[1];
    lv_total<1>(a.part + LV_P_RR * LV_SLOT, nb, rr);
    if (done) return;
    const int k = fs->k + 1;   // fs->k itself advances in k_fcg_true_fin
    const bool nan = !(rr[0] == rr[0]);
    const bool pending = !nan && (sqrt(rr[0]) < fs->atol || k >= fs->maxiter);
    if (blockIdx.x == 0 && threadIdx.x == 0) {   // fields no work-group reads here
        fs->rr = rr[0];
        fs->rho_old = fs->rho;
        fs->iters = k;
        fs->first = 0;
        fs->pending = pending ? 1 : 0;
        if (nan) {
            fs->done = 1;
            fs->info = -1;
        }
    }
    if (!pending) return;
    LV_CTX_LDS(a);
    const MGLevelDev L0 = a.mg->lev[0];
    const float w = a.mg->omega;
    double s[1] = {0.0};
    struct Tr {
        int li;
        double rt;
    };
    lv_cells_batch<4>(
        v.nx, L0.ny,
        [&](int i, int j) {
            const long c = ix(i, j, v.pc);
            Tr o;
            o.li = a.code[c] & 63;
            o.rt = v.b[c] -
                   apply_A(v.x, i, j, v.pc, v.nx, a.periodic, v.lut[o.li], v.ax, v.ay);
            return o;
        },
        [&](int i, int j, const Tr& o) {
            if (!o.li) return;
            const long c = ix(i, j, v.pc);
            v.r[c] = o.rt;
            s[0] += o.rt * o.rt;
            const float rf = (float)o.rt;
            L0.b[c] = rf;
            if (a.store_x0) L0.x[c] = w * rf / m.diag0[o.li];
        });
    lv_partial<1>(s, a.part + LV_P_PQ * LV_SLOT);
}

// one work-group: advance the iteration count; after a true residual, finish
// or restart the flexible CG from it
__global__ void __launch_bounds__(256) k_fcg_true_fin(LvArgs a, int nb, int net_used) {
    FcgState* fs = a.fs;
    if (fs->done) return;
    if (threadIdx.x == 0) {
        fs->k += 1;
        // a network that has not reduced ||r|| tenfold in LV_NET_CHECK
        // iterations is switched off for the context; the CG restarts
        if (net_used && !fs->net_off && !fs->pending && fs->k == LV_NET_CHECK &&
            !(fs->rr <= 1e-2 * fs->rr0)) {
            fs->net_off = 1;
            fs->first = 1;
        }
        fs->skip_net = fs->done | fs->net_off;
    }
    __syncthreads();
    if (!fs->pending) return;
    double rr[1];
    lv_total<1>(a.part + LV_P_PQ * LV_SLOT, nb, rr);
    if (threadIdx.x == 0) {
        fs->rr = rr[0];
        fs->reps += 1;
        fs->pending = 0;
        if (sqrt(rr[0]) < fs->atol) {
            fs->done = 1;
            fs->info = 0;
        } else if (!(rr[0] == rr[0])) {
            fs->done = 1;
            fs->info = -1;
        } else if (fs->k >= fs->maxiter) {
            fs->done = 1;
            fs->info = fs->maxiter;
        } else {
            fs->first = 1;
        }
        fs->skip_net = fs->done | fs->net_off;
    }
}

// ---- the coarse end of the V-cycle in one work-group ----------------------
// Levels Lc .. K of one preconditioner application for a 32 x 32 level K:
// the down-leg sweeps and restrictions from level Lc, the constant-vector
// statistics c and s, the network on the deflated, normalised level-K
// right-hand side (activations in LDS), the correction t_K = c + gain s y, the
// two level-K sweeps and the up-leg back to level Lc -- the work of
// k_lv_smooth / k_lv_restrict / k_lv_cmode / k_lv_net_in / the network /
// k_lv_net_out / k_lv_post on these levels, cell for cell the same FP32
// arithmetic, with __syncthreads between the sweeps instead of launches.
// One work-group of 1024 threads; returns at once after the solve finished.
struct CoarseNet {
    const unsigned short* wsw;   // lane-ordered matrix-core weights
    const float* w;              // FP32 parameters
    float gain;                  // sigma / D
    int use_net;
    int vcycle;                  // 1: V-cycle levels Lc .. coarsest, dense solve there
    double one_a_one;            // 1^T A_K 1
};

namespace {

// f over the nx x ny cells of a level by the whole work-group, U cells per
// thread at a time: gather (loads and arithmetic, no stores) for all U before
// scatter stores any, so that the loads of U cells are in flight together
template <int U, typename G, typename S>
__device__ __forceinline__ void wg_cells(int nx, int ny, G gather, S scatter) {
    const int n = nx * ny, bs = blockDim.x;
    using T = decltype(gather(1, 1));
    for (int q0 = threadIdx.x; q0 < n; q0 += U * bs) {
        T val[U];
        int ii[U], jj[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int q = q0 + u * bs < n ? q0 + u * bs : q0;
            ii[u] = q / ny + 1;
            jj[u] = q % ny + 1;
            val[u] = gather(ii[u], jj[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (q0 + u * bs < n) scatter(ii[u], jj[u], val[u]);
    }
    __syncthreads();
}

struct CellVal {
    float d, v;
};

// wg_cells with 4 cells per thread at a time on levels of >= 4 cells per
// thread, else 1 (the clamped extra gathers of a partial batch cost loads)
template <typename G, typename S>
__device__ __forceinline__ void wg_cells_auto(int nx, int ny, G gather, S scatter) {
    if (nx * ny >= 4 * (int)blockDim.x) wg_cells<4>(nx, ny, gather, scatter);
    else wg_cells<1>(nx, ny, gather, scatter);
}

// dst = src + w (b - A src) / d over level l, by the whole work-group
__device__ __forceinline__ void wg_smooth(const MGCtx& m, int l, const MGLevelDev& L, float w,
                                          const float* src, float* dst) {
    wg_cells_auto(
        L.nx, L.ny,
        [&](int i, int j) {
            const long c = ix(i, j, L.P);
            CellVal o;
            o.d = mg_diag(m, l, L, c);
            auto sv = [&](int p, int q) -> float { return src[ix(p, q, L.P)]; };
            o.v = src[c] + w * (L.b[c] - mg_apply(m, l, L, i, j, o.d, sv)) / o.d;
            return o;
        },
        [&](int i, int j, const CellVal& o) {
            if (o.d != 0.0f) dst[ix(i, j, L.P)] = o.v;
        });
}

// up-leg from level `from` to level Lc by the whole work-group: prolongation
// of the level below plus the first post-smoothing sweep, then the second
__device__ __forceinline__ void lv_up_leg(const MGCtx& m, const MGDev& mg, int from, int Lc) {
    const float w = mg.omega, wc = mg.wc;
    for (int l = from; l >= Lc; --l) {
        const MGLevelDev L = mg.lev[l];
        const MGLevelDev C = mg.lev[l + 1];
        auto corr = [&](int p, int q) -> float {   // loads in bounds on the ghost ring too
            const long c = ix(p, q, L.P);
            const float v = L.t[c] + wc * C.t[ix((p + 1) >> 1, (q + 1) >> 1, C.P)];
            return mg_diag(m, l, L, c) == 0.0f ? 0.0f : v;
        };
        wg_cells_auto(
            L.nx, L.ny,
            [&](int i, int j) {
                const long c = ix(i, j, L.P);
                CellVal o;
                o.d = mg_diag(m, l, L, c);
                const float y = corr(i, j);
                o.v = y + w * (L.b[c] - mg_apply(m, l, L, i, j, o.d, corr)) / o.d;
                return o;
            },
            [&](int i, int j, const CellVal& o) {
                if (o.d != 0.0f) L.x[ix(i, j, L.P)] = o.v;
            });
        wg_smooth(m, l, L, w, L.x, L.t);
    }
}

}  // namespace

__global__ void __launch_bounds__(1024) k_lv_coarse(LvArgs a, int Lc, int K, CoarseNet net) {
    FcgState* fs = a.fs;
    if (fs->done) return;
    __shared__ __align__(16) unsigned short S0[IMG_LDS_S0], W0a[IMG_LDS_S0];
    __shared__ __align__(16) unsigned short R[IMG_LDS_R];
    __shared__ double red[5 * 16];
    __shared__ float s_d0_[64];
    if (threadIdx.x < 64) s_d0_[threadIdx.x] = a.diag0[threadIdx.x];
    __syncthreads();
    const MGCtx m{a.mg, a.code, s_d0_, a.periodic, a.ax, a.ay};
    const MGDev& mg = *a.mg;
    const float w = mg.omega, wc = mg.wc;
    const bool use_net = net.use_net && !fs->net_off;
    // the last level of the down-leg: K, or the dense coarsest level
    const int Kd = net.vcycle ? mg.nlev - 1 : K;
    // down-leg: pre-smoothing sweep, restriction of the residual
    for (int l = Lc; l < Kd; ++l) {
        const MGLevelDev L = mg.lev[l];
        const MGLevelDev C = mg.lev[l + 1];
        wg_smooth(m, l, L, w, L.x, L.t);
        wg_cells_auto(
            C.nx, C.ny,
            [&](int ci, int cj) {
                float bc = 0.0f;
                for (int di = 0; di < 2; ++di)
                    for (int dj = 0; dj < 2; ++dj) {
                        const int i = 2 * ci - 1 + di, j = 2 * cj - 1 + dj;
                        const bool in = i <= L.nx && j <= L.ny;
                        const long c = ix(in ? i : 1, in ? j : 1, L.P);
                        const float d = mg_diag(m, l, L, c);
                        auto tv = [&](int p, int q) -> float { return L.t[ix(p, q, L.P)]; };
                        const float r = L.b[c] - mg_apply(m, l, L, in ? i : 1, in ? j : 1, d, tv);
                        if (in && d != 0.0f) bc += r;
                    }
                return bc;
            },
            [&](int ci, int cj, float bc) {
                const long cc = ix(ci, cj, C.P);
                C.b[cc] = bc;
                const float dc = C.diag[cc];
                C.x[cc] = dc != 0.0f ? w * bc / dc : 0.0f;
            });
    }
    if (net.vcycle) {
        // coarsest level: t = cinv b (cinv is zero in the columns of solid cells)
        const MGLevelDev C = mg.lev[Kd];
        const int n = mg.ncoarse;
        for (int k = threadIdx.x; k < n; k += blockDim.x) {
            float acc = 0.0f;
            for (int q = 0; q < n; ++q)
                acc += mg.cinv[(long)k * n + q] * C.b[ix(q / C.ny + 1, q % C.ny + 1, C.P)];
            C.t[ix(k / C.ny + 1, k % C.ny + 1, C.P)] = acc;
        }
        __syncthreads();
        lv_up_leg(m, mg, Kd - 1, Lc);
        return;
    }
    // level K: c = 1^T b / 1^T A 1 and s = RMS(b - c A 1) over the fluid cells
    const MGLevelDev LK = mg.lev[K];
    const int nK = LK.nx * LK.ny;
    double st[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k = threadIdx.x; k < nK; k += blockDim.x) {
        const int i = k / LK.ny + 1, j = k % LK.ny + 1;
        const long c = ix(i, j, LK.P);
        const float d = mg_diag(m, K, LK, c);
        if (d == 0.0f) continue;
        const double a1 = lv_a1(m, K, LK, i, j, d);
        const double bv = LK.b[c];
        st[0] += bv;
        st[1] += bv * bv;
        st[2] += bv * a1;
        st[3] += a1 * a1;
        st[4] += 1.0;
    }
    block_sum<5>(st, red);
    const double cm = net.one_a_one > 0.0 ? st[0] / net.one_a_one : 0.0;
    const double bb = st[1] - 2.0 * cm * st[2] + cm * cm * st[3];
    const double s = st[4] > 0.0 ? sqrt(fmax(bb, 0.0) / st[4]) : 0.0;
    if (threadIdx.x == 0) {
        fs->cm = cm;
        fs->s_in = s;
    }
    if (use_net) {
        // network input (FP32 in R), the network, t_K = c + gain s y
        float* xin = reinterpret_cast<float*>(R);
        for (int k = threadIdx.x; k < 1024; k += blockDim.x) {
            const int i = k / 32 + 1, j = k % 32 + 1;
            const long c = ix(i, j, LK.P);
            const float d = mg_diag(m, K, LK, c);
            float val = 0.0f;
            if (d != 0.0f && s > 0.0)
                val = (float)(((double)LK.b[c] - cm * (double)lv_a1(m, K, LK, i, j, d)) / s);
            xin[k] = val;
        }
        __syncthreads();
        const double gs = (double)net.gain * s;
        unet_img32(xin, S0, W0a, R, net.wsw, net.w, [&](int p, float y) {
            const int i = p / 32 + 1, j = p % 32 + 1;
            const long c = ix(i, j, LK.P);
            if (mg_diag(m, K, LK, c) != 0.0f) LK.t[c] = (float)(cm + gs * (double)y);
        });
    } else {
        for (int k = threadIdx.x; k < nK; k += blockDim.x) {
            const int i = k / LK.ny + 1, j = k % LK.ny + 1;
            const long c = ix(i, j, LK.P);
            if (mg_diag(m, K, LK, c) != 0.0f) LK.t[c] = (float)cm;
        }
    }
    __syncthreads();
    wg_smooth(m, K, LK, w, LK.t, LK.x);
    wg_smooth(m, K, LK, w, LK.x, LK.t);
    lv_up_leg(m, mg, K - 1, Lc);
}

// ---- the complete V-cycle's coarse end in LDS ------------------------------
// Levels Lc .. coarsest (Lc: the first level of at most 64 x 64 cells) of one
// V-cycle application, held in LDS for the whole launch: per level b, x, t and
// the coefficients diag, cE, cN at the tight pitch ny + 2 with the zero ghost
// ring.  The sweeps, restrictions, dense coarsest solve and prolongations are
// the expressions of k_lv_smooth / k_lv_restrict / k_lv_dense / k_lv_post in
// the same order, so the result is bit-identical to the launch sequence; only
// level Lc's b comes from global memory (the previous launch's restriction)
// and only its t goes back (for the next launch's prolongation).
constexpr int VC_POOL = 36864;           // LDS floats (144 KB)
constexpr int VC_ARR = 6;                // b, x, t, diag, cE, cN per level
constexpr int VC_MAXL = 8;               // levels held in LDS
constexpr int VC_MAXC = 64;              // cells of the dense coarsest level

// LDS floats needed for levels Lc .. nlev-1 (host and device; with the
// float4 alignment of the two parts of the pool, see vc_assign)
__host__ __device__ inline long vc_need(const MGDev& mg, int Lc) {
    long s = 0;
    for (int l = Lc; l < mg.nlev; ++l)
        s += (long)VC_ARR * (mg.lev[l].nx + 2) * (mg.lev[l].ny + 2);
    return s + 8;
}

// one LDS level: sizes, the offsets of its coefficients (diag, cE, cN) and
// of its vectors (b, x, t) in the pool, the
// multiplier of the exact division q / ny for q < 2^16 (m = ceil(2^32 / ny)),
// and the global arrays it is loaded from / written back to (the right-hand
// side and correction of instance blockIdx.x at gb / gt + blockIdx.x * istride;
// the coefficients are shared)
struct VcLev {
    int nx, ny, pw, sz, coff, voff, P;
    unsigned long long m;
    long istride;
    const float *gd, *ge, *gn, *gb;
    float* gt;
};

// arguments of k_lv_vcoarse, built on the host from the hierarchy (kernel
// arguments: no descriptor loads on the device).  One work-group per instance
// (the single case: one); an instance whose solve has finished (the int at
// done + blockIdx.x * done_stride bytes) writes nothing back.
struct VcArgs {
    const char* done;
    long done_stride;
    const float* cinv;       // dense inverse of the coarsest level (shared)
    const float4* image;     // the coefficient part of the pool as it starts (shared)
    int image_n4;            // its length in float4
    int zero_off4, zero_n4;  // the vector part (float4 offset, length): zero at the start
    float omega, wc;
    int periodic, nl, nc;    // levels in LDS, cells of the coarsest level
    int total;               // sum of sz over the levels
    VcLev lev[VC_MAXL];
};

// The pool's layout for levels with nx, ny, pw, sz set: the coefficients of
// every level first (the part the launch copies from the image), then from
// a float4 boundary the vectors b, x, t of every level (the part the launch
// zeroes).  Returns the floats used (at most vc_need).
inline long vc_assign(VcArgs& v) {
    long c = 0;
    for (int q = 0; q < v.nl; ++q) {
        v.lev[q].coff = (int)c;
        c += 3L * v.lev[q].sz;
    }
    const long vb = (c + 3) / 4 * 4;
    long x = vb;
    for (int q = 0; q < v.nl; ++q) {
        v.lev[q].voff = (int)x;
        x += 3L * v.lev[q].sz;
    }
    v.image_n4 = (int)((c + 3) / 4);
    v.zero_off4 = (int)(vb / 4);
    v.zero_n4 = (int)((x - vb + 3) / 4);
    return vb + 4L * v.zero_n4;
}

// The image of the coefficient part of the pool, built on the host from the
// levels' device arrays (pitch P): per level diag, cE, cN at the tight pitch
// ny + 2 with their ghost rings, at the offsets of v.lev (vc_assign).
inline int vc_build_image(const VcArgs& v, std::vector<float>& img) {
    img.assign((size_t)v.image_n4 * 4, 0.0f);
    std::vector<float> h;
    for (int q = 0; q < v.nl; ++q) {
        const VcLev& L = v.lev[q];
        const size_t cnt = (size_t)(L.nx + 2) * L.P;
        h.resize(cnt);
        const float* src[3] = {L.gd, L.ge, L.gn};
        for (int a = 0; a < 3; ++a) {
            if (hipMemcpy(h.data(), src[a], cnt * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
                return -1;
            float* dst = img.data() + L.coff + a * L.sz;
            for (int i = 0; i <= L.nx + 1; ++i)
                for (int j = 0; j <= L.ny + 1; ++j) dst[i * L.pw + j] = h[(size_t)i * L.P + j];
        }
    }
    return 0;
}

namespace {

struct VcView {
    int nx, ny, pw;
    unsigned long long m;
    float *b, *x, *t, *d, *e, *n;
};

__device__ __forceinline__ VcView vc_view(float* pool, const VcLev& L) {
    VcView v;
    v.nx = L.nx;
    v.ny = L.ny;
    v.pw = L.pw;
    v.m = L.m;
    float* p = pool + L.voff;
    v.b = p;
    v.x = p + L.sz;
    v.t = p + 2 * L.sz;
    float* c = pool + L.coff;
    v.d = c;
    v.e = c + L.sz;
    v.n = c + 2 * L.sz;
    return v;
}

// interior cell q (row-major, 0-based) -> (i, j), 1-based
__device__ __forceinline__ void vc_ij(const VcView& L, int q, int& i, int& j) {
    const int r = (int)(((unsigned long long)(unsigned)q * L.m) >> 32);
    i = r + 1;
    j = q - r * L.ny + 1;
}

// (A v)(i, j) on an LDS level, mg_apply's expression for levels >= 1
__device__ __forceinline__ float vc_apply(const VcView& L, int periodic, int i, int j, float dc,
                                          const float* v) {
    const int iE = (periodic && i == L.nx) ? 1 : i + 1;
    const int iW = (periodic && i == 1) ? L.nx : i - 1;
    const int k = i * L.pw + j;
    return dc * v[k] - L.e[k] * v[iE * L.pw + j] - L.e[iW * L.pw + j] * v[iW * L.pw + j] -
           L.n[k] * v[k + 1] - L.n[k - 1] * v[k - 1];
}

// f over cells 0..n-1 of an LDS level by the whole work-group, U per thread
// at a time: gather(q) (LDS loads and arithmetic, no stores) for all U before
// scatter(q, value) stores any, so that the U cells' LDS loads are in flight
// together instead of one cell's dependent chain after the other
template <int U, typename G, typename S>
__device__ __forceinline__ void vc_batch(int n, G gather, S scatter) {
    const int bs = blockDim.x;
    using T = decltype(gather(0));
    for (int q0 = threadIdx.x; q0 < n; q0 += U * bs) {
        T val[U];
#pragma unroll
        for (int u = 0; u < U; ++u) val[u] = gather(q0 + u * bs < n ? q0 + u * bs : q0);
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (q0 + u * bs < n) scatter(val[u]);
    }
}

// vc_batch with 4 cells per thread at a time on levels of at least 4 cells
// per thread, else 1
template <typename G, typename S>
__device__ __forceinline__ void vc_cells(int n, G gather, S scatter) {
    if (n >= 4 * (int)blockDim.x) vc_batch<4>(n, gather, scatter);
    else vc_batch<1>(n, gather, scatter);
}

// a cell's LDS index, diagonal (0: not fluid) and new value
struct VcCell {
    int k;
    float d, v;
};

// dst = src + w (b - A src) / d on the fluid cells of an LDS level
__device__ __forceinline__ void vc_smooth(const VcView& L, int periodic, float w, const float* src,
                                          float* dst) {
    vc_cells(
        L.nx * L.ny,
        [&](int q) {
            int i, j;
            vc_ij(L, q, i, j);
            VcCell o;
            o.k = i * L.pw + j;
            o.d = L.d[o.k];
            o.v = src[o.k] + w * (L.b[o.k] - vc_apply(L, periodic, i, j, o.d, src)) / o.d;
            return o;
        },
        [&](const VcCell& o) {
            if (o.d != 0.0f) dst[o.k] = o.v;
        });
    __syncthreads();
}

}  // namespace

__global__ void __launch_bounds__(1024) k_lv_vcoarse(VcArgs a) {
    __shared__ __align__(16) float pool[VC_POOL];
    __shared__ float s_bc[VC_MAXC];
    const int nl = a.nl, nc = a.nc, per = a.periodic;
    const float w = a.omega, wc = a.wc;
    const int tid = threadIdx.x, bs = blockDim.x;
    // every global load of the launch in flight at once: the solve's state,
    // this lane's row of the dense inverse (wave 0), the pool's image and
    // level Lc's right-hand side.  (An instance whose solve has finished
    // still runs through the LDS work, which needs no global memory, so that
    // no wave waits for that flag's load; it writes nothing back.)
    const long inst = blockIdx.x;
    const int done = *reinterpret_cast<const int*>(a.done + inst * a.done_stride);
    float crow[VC_MAXC];
    if (tid < nc) {
#pragma unroll
        for (int q = 0; q < VC_MAXC; ++q) crow[q] = q < nc ? a.cinv[(long)tid * nc + q] : 0.0f;
    }
    // the coefficient part is at most half of the pool
    constexpr int NI = (VC_POOL / 8 + 1 + 1023) / 1024, NZ = NI, NB = (64 * 64 + 1023) / 1024;
    float4 img[NI];
#pragma unroll
    for (int u = 0; u < NI; ++u) {
        const int k = tid + u * bs;
        img[u] = k < a.image_n4 ? a.image[k] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    const VcLev& G = a.lev[0];
    const VcView L0 = vc_view(pool, G);
    float rb[NB];
    int kb[NB];
#pragma unroll
    for (int u = 0; u < NB; ++u) {
        const int c = tid + u * bs;
        kb[u] = -1;
        rb[u] = 0.0f;
        if (c >= L0.nx * L0.ny) continue;
        int i, j;
        vc_ij(L0, c, i, j);
        rb[u] = G.gb[inst * G.istride + ix(i, j, G.P)];
        kb[u] = i * L0.pw + j;
    }
#pragma unroll
    for (int u = 0; u < NI; ++u) {
        const int k = tid + u * bs;
        if (k < a.image_n4) reinterpret_cast<float4*>(pool)[k] = img[u];
    }
#pragma unroll
    for (int u = 0; u < NZ; ++u) {
        const int k = tid + u * bs;
        if (k < a.zero_n4)
            reinterpret_cast<float4*>(pool)[a.zero_off4 + k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    __syncthreads();
    // level Lc: b and the first sweep from zero, x = w b / d
#pragma unroll
    for (int u = 0; u < NB; ++u) {
        if (kb[u] < 0) continue;
        const float d = L0.d[kb[u]];
        L0.b[kb[u]] = rb[u];
        L0.x[kb[u]] = d != 0.0f ? w * rb[u] / d : 0.0f;
    }
    __syncthreads();
    // down-leg: second pre-smoothing sweep, residual, restriction
    for (int q = 0; q < nl - 1; ++q) {
        const VcView L = vc_view(pool, a.lev[q]);
        const VcView C = vc_view(pool, a.lev[q + 1]);
        vc_smooth(L, per, w, L.x, L.t);
        // the residual b - A t of the aggregate's fluid cells, summed (the
        // expression and order of k_lv_restrict)
        const bool last = q + 1 == nl - 1;   // (the coarsest level's x is not used)
        vc_cells(
            C.nx * C.ny,
            [&](int c) {
                int ci, cj;
                vc_ij(C, c, ci, cj);
                VcCell o;
                o.k = ci * C.pw + cj;
                float bc = 0.0f;
                for (int di = 0; di < 2; ++di)
                    for (int dj = 0; dj < 2; ++dj) {
                        const int i = 2 * ci - 1 + di, j = 2 * cj - 1 + dj;
                        if (i > L.nx || j > L.ny) continue;
                        const int k = i * L.pw + j;
                        const float d = L.d[k];
                        if (d != 0.0f) bc += L.b[k] - vc_apply(L, per, i, j, d, L.t);
                    }
                o.v = bc;
                o.d = last ? (float)c : C.d[o.k];   // the last level: the compact index
                return o;
            },
            [&](const VcCell& o) {
                C.b[o.k] = o.v;
                if (!last) {
                    C.x[o.k] = o.d != 0.0f ? w * o.v / o.d : 0.0f;
                } else {
                    const int c = (int)o.d;
                    if (c < VC_MAXC) s_bc[c] = o.v;   // the coarsest right-hand side, compact
                }
            });
        __syncthreads();
    }
    // coarsest level: t = cinv b, the sum in k_lv_dense's order
    {
        const VcView C = vc_view(pool, a.lev[nl - 1]);
        if (tid < nc) {
            float acc = 0.0f;
#pragma unroll
            for (int q = 0; q < VC_MAXC; ++q)
                if (q < nc) acc += crow[q] * s_bc[q];
            int i, j;
            vc_ij(C, tid, i, j);
            C.t[i * C.pw + j] = acc;
        }
        __syncthreads();
    }
    // up-leg: prolongation + first post-smoothing sweep into x, second into t
    for (int q = nl - 2; q >= 0; --q) {
        const VcView L = vc_view(pool, a.lev[q]);
        const VcView C = vc_view(pool, a.lev[q + 1]);
        // y = t + wc t_(l+1) in place on the fluid cells (0 stays on the
        // others, as k_lv_post's corr has it), then the two sweeps
        vc_cells(
            L.nx * L.ny,
            [&](int c) {
                int i, j;
                vc_ij(L, c, i, j);
                VcCell o;
                o.k = i * L.pw + j;
                o.d = L.d[o.k];
                o.v = L.t[o.k] + wc * C.t[((i + 1) >> 1) * C.pw + ((j + 1) >> 1)];
                return o;
            },
            [&](const VcCell& o) {
                if (o.d != 0.0f) L.t[o.k] = o.v;
            });
        __syncthreads();
        vc_smooth(L, per, w, L.t, L.x);
        vc_smooth(L, per, w, L.x, L.t);
    }
    // level Lc's t for the prolongation launch of level Lc - 1
    {
        const VcLev& G = a.lev[0];
        const VcView L = vc_view(pool, G);
        const int n = L.nx * L.ny;
        for (int c = tid; c < n; c += bs) {
            int i, j;
            vc_ij(L, c, i, j);
            const int k = i * L.pw + j;
            if (!done && L.d[k] != 0.0f) G.gt[inst * G.istride + ix(i, j, G.P)] = L.t[k];
        }
    }
}

// initial iterate of a time step's solve: x <- x + theta (x - x_prev) with
// theta = dt / dt_old, and x_prev <- the old x (hist = 0: the copy only); a
// step after the end of the run (st->done) leaves both as they are
__global__ void __launch_bounds__(256) k_x0_extrap(double* __restrict__ x, double* __restrict__ xprev,
                                                   long n, const DevState* st, int hist) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n || st->done) return;
    const double xn = x[k];
    if (hist) {
        const double th = st->dt_old > 0.0 ? st->dt / st->dt_old : 1.0;
        x[k] = xn + th * (xn - xprev[k]);
    }
    xprev[k] = xn;
}

// ---- start and end of a learned projection as band launches ---------------
// k_project's steps before and after the solve, for the learned solver: the
// same cell arithmetic (right-hand side, compatibility shift, initial residual and
// first V-cycle input; pressure mean, write-back, velocity correction and
// residual monitor), with the sums taken per work-group and added by the
// consuming kernel in work-group order instead of grid barriers.  Band
// kernels: min(nx, LV_MAXB) work-groups of 512 threads.

namespace {

// rows of logical band b out of nb over rows 1..nx
__device__ __forceinline__ void lp_band(int nx, int b, int nb, int& i0, int& i1) {
    i0 = 1 + (int)(((long)nx * b) / nb);
    i1 = 1 + (int)(((long)nx * (b + 1)) / nb);
}

// the end launches of a learned projection are enqueued after every batch of
// iterations and act once: when the solve has finished successfully (or, with
// end_always, after a failed one that is not continued by the fallback)
__device__ __forceinline__ bool lp_end_idle(const ProjArgs& a) {
    if (a.st->done) return true;
    if (a.end_always) return false;
    return !(a.fcg->done && a.fcg->info == 0);
}

__device__ __forceinline__ void lp_means(const ProjArgs& a, const double* slot, int nb,
                                         double (&mean)[CFD_MAX_COMP]) {
    double cs[CFD_MAX_COMP];
    lv_total<CFD_MAX_COMP>(slot, nb, cs);
#pragma unroll
    for (int k = 0; k < CFD_MAX_COMP; ++k)
        mean[k] = (a.ncomp > 0 && k < a.ncomp) ? cs[k] / a.comp_count[k] : 0.0;
}

}  // namespace

// b = div(F, G) / dt (or the given right-hand side) on the fluid cells, and
// the partial sums per Neumann component; with x0mode the initial iterate as
// k_x0_extrap forms it (1: x_prev <- x, 2: also x <- x + theta (x - x_prev)),
// on the interior cells (x and x_prev are zero on the ghost ring)
__global__ void __launch_bounds__(512) k_lp_rhs(ProjArgs a, Geo g, double* part,
                                                double* __restrict__ xprev, int x0mode) {
    const DevState* st = a.st;
    if (st->done) return;
    const double dt = a.dt_override > 0.0 ? a.dt_override : st->dt;
    const double th = st->dt_old > 0.0 ? st->dt / st->dt_old : 1.0;
    double cs[CFD_MAX_COMP];
#pragma unroll
    for (int k = 0; k < CFD_MAX_COMP; ++k) cs[k] = 0.0;
    struct Rhs {
        int fluid, cm;
        double rhs, x, xp;
    };
    lv_cells_batch<4>(
        g.nx, g.ny,
        [&](int i, int j) {
            const long c = ix(i, j, g.pc);
            Rhs o;
            o.fluid = a.code[c] & 63;
            o.cm = a.comp[c];
            if (a.rhs_in) {
                o.rhs = a.rhs_in[c];
            } else {
                const double div = (a.F[ix(i, j, g.pu)] - a.F[ix(i - 1, j, g.pu)]) / g.dx +
                                   (a.G[ix(i, j, g.pv)] - a.G[ix(i, j - 1, g.pv)]) / g.dy;
                o.rhs = div / dt;
            }
            o.x = x0mode ? a.x[c] : 0.0;
            o.xp = x0mode == 2 ? xprev[c] : 0.0;
            return o;
        },
        [&](int i, int j, const Rhs& o) {
            if (x0mode) {
                const long c = ix(i, j, g.pc);
                if (x0mode == 2) a.x[c] = o.x + th * (o.x - o.xp);
                xprev[c] = o.x;
            }
            if (!o.fluid) return;
            a.b[ix(i, j, g.pc)] = o.rhs;
#pragma unroll
            for (int k = 0; k < CFD_MAX_COMP; ++k) cs[k] += (o.cm == k + 1) ? o.rhs : 0.0;
        });
    lv_partial<CFD_MAX_COMP>(cs, part + LV_P_CM * LV_SLOT);
}

// b = -(rhs - mean), r = b - A x, the first V-cycle input (lev[0].b = r,
// lev[0].x = w r / d), partial b.b and r.r; nb = work-groups of k_lp_rhs
__global__ void __launch_bounds__(512) k_lp_res(ProjArgs a, Geo g, LvArgs la, int nb) {
    if (a.st->done) return;
    double mean[CFD_MAX_COMP];
    lp_means(a, la.part + LV_P_CM * LV_SLOT, nb, mean);
    __shared__ double s_diag[64];
    __shared__ float s_diagf[64];
    if (threadIdx.x < 64) {
        s_diag[threadIdx.x] = a.lut[threadIdx.x];
        s_diagf[threadIdx.x] = (float)a.lut[threadIdx.x];
    }
    __syncthreads();
    const double ax = 1.0 / (g.dx * g.dx), ay = 1.0 / (g.dy * g.dy);
    const MGLevelDev L0 = la.mg->lev[0];
    const float w = la.mg->omega;
    double s2[2] = {0.0, 0.0};
    struct Res {
        int li;
        double bv, rv;
    };
    lv_cells_batch<4>(
        g.nx, g.ny,
        [&](int i, int j) {
            const long c = ix(i, j, g.pc);
            Res o;
            o.li = a.code[c] & 63;
            double bv = a.b[c];
            const int cm = a.comp[c];
            if (cm) bv = bv - sel_comp(mean, cm);
            o.bv = -bv;
            o.rv = o.bv - apply_A(a.x, i, j, g.pc, g.nx, g.periodic, s_diag[o.li], ax, ay);
            return o;
        },
        [&](int i, int j, const Res& o) {
            if (!o.li) return;
            const long c = ix(i, j, g.pc);
            a.b[c] = o.bv;
            a.r[c] = o.rv;
            s2[0] += o.bv * o.bv;
            s2[1] += o.rv * o.rv;
            const float rf = (float)o.rv;
            L0.b[c] = rf;
            if (la.store_x0) L0.x[c] = w * rf / s_diagf[o.li];
        });
    lv_partial<2>(s2, la.part + LV_P_RR * LV_SLOT);
}

// one work-group: ||b||, the flexible CG's initial state (a zero right-hand
// side finishes the solve: x is zeroed by k_lp_xsum)
__global__ void __launch_bounds__(256) k_lp_init(ProjArgs a, LvArgs la, int nb) {
    if (a.st->done) return;
    double s2[2];
    lv_total<2>(la.part + LV_P_RR * LV_SLOT, nb, s2);
    if (threadIdx.x != 0) return;
    FcgState* fs = a.fcg;
    const double bnorm = sqrt(s2[0]);
    const double atol = a.rtol * bnorm;
    const double rr = s2[1];
    fs->bnorm = bnorm;
    fs->atol = atol;
    fs->rr = bnorm == 0.0 ? 0.0 : rr;
    fs->rho_old = 1.0;
    fs->k = 0;
    fs->first = 1;
    fs->pending = 0;
    fs->iters = 0;
    fs->info = 0;
    fs->reps = 0;
    fs->maxiter = a.maxiter;
    fs->done = (bnorm == 0.0 || sqrt(rr) < atol) ? 1 : 0;
    fs->rr0 = fs->rr;
    fs->skip_net = fs->done | fs->net_off;
}

// partial sums of x per Neumann component (x := 0 first on the fluid cells
// when the right-hand side was zero: SciPy returns b itself)
__global__ void __launch_bounds__(512) k_lp_xsum(ProjArgs a, Geo g, double* part) {
    if (lp_end_idle(a)) return;
    const bool zero = a.fcg->bnorm == 0.0;
    double cs[CFD_MAX_COMP];
#pragma unroll
    for (int k = 0; k < CFD_MAX_COMP; ++k) cs[k] = 0.0;
    struct Xs {
        double x;
        int cm, fluid;
    };
    lv_cells_batch<4>(
        g.nx, g.ny,
        [&](int i, int j) {
            const long c = ix(i, j, g.pc);
            Xs o;
            o.x = a.x[c];
            o.cm = a.comp[c];
            o.fluid = a.code[c] & 63;
            return o;
        },
        [&](int i, int j, const Xs& o) {
            double xv = o.x;
            if (zero && o.fluid) {
                a.x[ix(i, j, g.pc)] = 0.0;
                xv = 0.0;
            }
#pragma unroll
            for (int k = 0; k < CFD_MAX_COMP; ++k) cs[k] += (o.cm == k + 1) ? xv : 0.0;
        });
    lv_partial<CFD_MAX_COMP>(cs, part + LV_P_CM * LV_SLOT);
}

// p = x - mean on the fluid cells, 0 elsewhere and on the ghost columns; the
// ghost rows 0 and nx + 1 (periodic copies of cells) by the first and last
// band.  With CORRECT also u = F - dt grad p, v = G - dt grad p (boundary faces
// as in k_project) and the partial max |u_new - u_old| over the faces between
// fluid cells.  The pressure of the neighbouring cells is formed from x here
// too (the same expression as the stored p), so that no work-group waits for
// another band's p.  Thread cells (i, j) of the band take the faces u(i, j)
// and v(i, j) (4 at a time, their loads in flight together); the bottom faces
// v(i, 0) and the left faces u(0, j) are the band's and the first band's
// extra work.
template <bool CORRECT>
__global__ void __launch_bounds__(512) k_lp_end(ProjArgs a, Geo g, LvArgs la, int nb) {
    if (lp_end_idle(a)) return;
    double mean[CFD_MAX_COMP];
    lp_means(a, la.part + LV_P_CM * LV_SLOT, nb, mean);
    const int nx = g.nx, ny = g.ny, pc = g.pc;
    const int periodic = g.periodic;
    auto pval = [&](int i, int j) -> double {   // loads and selects, no branches
        const long c = ix(i, j, pc);
        const int fl = a.code[c] & 63;
        const int cm = a.comp[c];
        const double xv = a.x[c];
        const double v = cm ? xv - sel_comp(mean, cm) : xv;
        return fl ? v : 0.0;
    };
    auto fluid = [&](int i, int j) -> bool { return a.code[ix(i, j, pc)] & 63; };
    const int b = xcd_block(blockIdx.x, gridDim.x);
    int i0, i1;
    lp_band(nx, b, gridDim.x, i0, i1);
    if (!CORRECT) {
        lv_cells_batch<4>(
            nx, ny, [&](int i, int j) { return pval(i, j); },
            [&](int i, int j, double v) { a.p[ix(i, j, pc)] = v; });
    } else {
        const double dt = a.dt_override > 0.0 ? a.dt_override : a.st->dt;
        double res[1] = {0.0};
        struct Faces {
            double p, un, vn, du, dv;
            bool uf, vf, vset;
        };
        lv_cells_batch<4>(
            nx, ny,
            [&](int i, int j) {
                Faces o;
                const long cu = ix(i, j, g.pu), cv = ix(i, j, g.pv);
                const double F = a.F[cu], G = a.G[cv];
                o.p = pval(i, j);
                o.uf = false;
                if (i < nx) {
                    o.uf = fluid(i, j) && fluid(i + 1, j);
                    o.un = o.uf ? F - dt * (pval(i + 1, j) - o.p) / g.dx : F;
                } else if (periodic) {
                    o.uf = fluid(nx, j) && fluid(1, j);
                    o.un = o.uf ? F - dt * (pval(1, j) - o.p) / g.dx : F;
                } else {
                    o.un = (a.code[ix(nx, j, pc)] & 64) ? F + 2 * dt * o.p / g.dx : F;
                }
                o.du = o.uf ? fabs(o.un - a.u[cu]) : 0.0;
                o.vf = false;
                o.vset = true;
                if (j < ny) {
                    o.vf = fluid(i, j) && fluid(i, j + 1);
                    o.vn = o.vf ? G - dt * (pval(i, j + 1) - o.p) / g.dy : G;
                } else if (periodic) {
                    o.vset = false;   // solver.py:300-306 returns before the boundary faces
                    o.vn = G;
                } else {
                    o.vn = (a.code[ix(i, ny, pc)] & 128) ? G + 2 * dt * o.p / g.dy : G;
                }
                o.dv = o.vf ? fabs(o.vn - a.v[cv]) : 0.0;
                return o;
            },
            [&](int i, int j, const Faces& o) {
                a.p[ix(i, j, pc)] = o.p;
                if (o.uf) res[0] = nanmax(res[0], o.du);
                a.u[ix(i, j, g.pu)] = o.un;
                if (periodic && i == nx) a.u[ix(0, j, g.pu)] = o.un;
                if (o.vf) res[0] = nanmax(res[0], o.dv);
                if (o.vset) a.v[ix(i, j, g.pv)] = o.vn;
            });
        if (!periodic) {
            for (int i = i0 + (int)threadIdx.x; i < i1; i += blockDim.x) {   // v(i, 0)
                const long cv = ix(i, 0, g.pv);
                const double G = a.G[cv];
                a.v[cv] = (a.code[ix(i, 1, pc)] & 128) ? G - 2 * dt * pval(i, 1) / g.dy : G;
            }
        }
        if (b == 0 && !periodic) {
            for (int j = 1 + threadIdx.x; j <= ny; j += blockDim.x) {
                const long cu = ix(0, j, g.pu);
                const double F = a.F[cu];
                a.u[cu] = (a.code[ix(1, j, pc)] & 64) ? F - 2 * dt * pval(1, j) / g.dx : F;
            }
        }
        __shared__ double lds[16];
        block_max<1>(res, lds);
        if (threadIdx.x == 0) la.part[LV_P_PQ * LV_SLOT + blockIdx.x] = res[0];
    }
    for (int i = i0 + (int)threadIdx.x; i < i1; i += blockDim.x) {   // the ghost columns
        a.p[ix(i, 0, pc)] = 0.0;
        a.p[ix(i, ny + 1, pc)] = 0.0;
    }
    if (b == 0)
        for (int j = threadIdx.x; j <= ny + 1; j += blockDim.x)
            a.p[ix(0, j, pc)] = (periodic && j >= 1 && j <= ny) ? pval(nx, j) : 0.0;
    if (b == (int)gridDim.x - 1)
        for (int j = threadIdx.x; j <= ny + 1; j += blockDim.x)
            a.p[ix(nx + 1, j, pc)] = (periodic && j >= 1 && j <= ny) ? pval(1, j) : 0.0;
}
template __global__ void k_lp_end<false>(ProjArgs, Geo, LvArgs, int);
template __global__ void k_lp_end<true>(ProjArgs, Geo, LvArgs, int);

// one work-group: the solve's statistics into the step state, the residual
// monitor, the periodic corner faces; then (finished or not) the flexible
// CG's state and the sequence number `seq` into the host's poll record
__global__ void __launch_bounds__(256) k_lp_fin(ProjArgs a, Geo g, double* part, int nb,
                                                LpPoll* poll, unsigned seq) {
    DevState* st = a.st;
    if (!lp_end_idle(a)) {
        double res = 0.0;
        if (a.correct)
            for (int k = threadIdx.x; k < nb; k += blockDim.x)
                res = nanmax(res, part[LV_P_PQ * LV_SLOT + k]);
        __shared__ double lds[16];
        double m[1] = {res};
        block_max<1>(m, lds);
        if (threadIdx.x == 0) {
            const FcgState* fs = a.fcg;
            st->cg_iters = fs->iters;
            st->cg_info = fs->info;
            st->cg_outer = fs->reps;
            st->cg_bnorm = fs->bnorm;
            st->cg_relres = fs->bnorm > 0.0 ? sqrt(fs->rr) / fs->bnorm : 0.0;
            if (a.correct) {
                if (a.dt_override <= 0.0) st->residual = m[0];
                if (g.periodic) {
                    a.u[ix(0, 0, g.pu)] = a.u[ix(g.nx, 0, g.pu)];
                    a.u[ix(0, g.ny + 1, g.pu)] = a.u[ix(g.nx, g.ny + 1, g.pu)];
                }
            }
        }
    }
    if (poll && threadIdx.x == 0) {
        poll->s = *a.fcg;
        __threadfence_system();
        __hip_atomic_store(&poll->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// p := 0 on the whole padded array (after the initial projection)
__global__ void __launch_bounds__(256) k_lp_pzero(ProjArgs a, long n) {
    if (lp_end_idle(a)) return;
    double* p = a.p;
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0.0;
}
