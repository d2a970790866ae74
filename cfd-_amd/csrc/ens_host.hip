// Host side of the batched cavity ensemble (ens.hip): classes of equal-size
// instances, one launch sequence per step for all of them, the learned
// flexible CG with one batched network pass (or the LDS coarse end) per
// iteration, and the predictor tile shape chosen per class by timing the
// candidates at setup.  With graphs a step is one graph launch: the
// pre-solve kernels, `cap` solver iterations, k_e_cap, the post-solve kernels
// and k_e_flip, all controlled on the device (EnsCtl); the host checks the
// completed steps once per cfd_ens_step call and adapts `cap` to the
// iterations the steps took.

static_assert(ENS_MAXLEV == CFD_ENS_MAX_LEVELS, "level tables of the tuning record");

struct cfd_ens {
    int ncls = 0;
    EnsClass h[CFD_ENS_MAX_CLASSES];
    EnsClass* d = nullptr;              // device copies
    int pred_tx[CFD_ENS_MAX_CLASSES], pred_ty[CFD_ENS_MAX_CLASSES];
    float pred_us[CFD_ENS_MAX_CLASSES];
    // work-group shape of the pointwise V-cycle kernels per class and level,
    // and the FP64 vector kernels' cells per thread and batch per class
    int sw_tx[CFD_ENS_MAX_CLASSES][ENS_MAXLEV], sw_ty[CFD_ENS_MAX_CLASSES][ENS_MAXLEV];
    float sw_us[CFD_ENS_MAX_CLASSES][ENS_MAXLEV], sw_us0[CFD_ENS_MAX_CLASSES][ENS_MAXLEV];
    int vec_u[CFD_ENS_MAX_CLASSES];
    float vec_us[CFD_ENS_MAX_CLASSES], vec_us0[CFD_ENS_MAX_CLASSES];
    int nimg = 0;
    float* img = nullptr;               // network input batch (nimg x 32 x 32)
    cfd_unet* net = nullptr;
    int use_net = 0;
    int* notdone = nullptr;
    int* notdone_h = nullptr;           // pinned
    EnsCtl* ctl = nullptr;              // device step control (ens.hip)
    int cap = 16;                       // solver iterations in the step graph
    int cap_graph = 0;                  // cap of the captured graph (0: none)
    long long steps = 0;
    int prev_iters = 0;
    int use_graph = 0;
    int x0_extrap = 0;                  // extrapolated initial iterates
    int vcycle = 0;                     // coarse end: 1 the V-cycle to the dense solve in LDS
    int vlc[CFD_ENS_MAX_CLASSES];       // first level of the LDS coarse end per class
    VcArgs vargs[CFD_ENS_MAX_CLASSES];
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_x = nullptr;
    hipStream_t stream = nullptr;
    std::vector<void*> allocs;
};

namespace {

template <typename T>
int ealloc(cfd_ens* e, T** p, size_t n) {
    void* q = nullptr;
    HIPC(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
    HIPC(hipMemset(q, 0, std::max<size_t>(n, 1) * sizeof(T)));
    HIPC(hipDeviceSynchronize());   // the fill is done before any stream uses the array
    e->allocs.push_back(q);
    *p = (T*)q;
    return 0;
}

#define E_LAUNCH(kern, grid, block, ...)                                              \
    do {                                                                              \
        hipLaunchKernelGGL(kern, grid, block, 0, e->stream, __VA_ARGS__);             \
        HIPC(hipGetLastError());                                                      \
    } while (0)

// grid and work-group of the pointwise V-cycle kernels on level l of class k
// (tiles of sw_ty rows x sw_tx columns, y = instance)
dim3 sw_grid(const cfd_ens* e, int k, int l) {
    const int n = e->h[k].lev[l].n, tx = e->sw_tx[k][l], ty = e->sw_ty[k][l];
    return dim3((unsigned)(((n + ty - 1) / ty) * ((n + tx - 1) / tx)), (unsigned)e->h[k].B);
}
dim3 sw_block(const cfd_ens* e, int k, int l) { return dim3(e->sw_tx[k][l], e->sw_ty[k][l]); }

// the FP64 vector kernels with the class's cells per thread and batch
#define E_LAUNCH_VEC(kern, u, grid, block, ...)                       \
    do {                                                              \
        switch (u) {                                                  \
        case 2: E_LAUNCH(kern<2>, grid, block, __VA_ARGS__); break;   \
        case 4: E_LAUNCH(kern<4>, grid, block, __VA_ARGS__); break;   \
        case 16: E_LAUNCH(kern<16>, grid, block, __VA_ARGS__); break; \
        default: E_LAUNCH(kern<8>, grid, block, __VA_ARGS__); break;  \
        }                                                             \
    } while (0)

// the fine level's last sweep with the dot products and the flexible CG's
// vector kernels of class k
int ens_fcg_tail(cfd_ens* e, int k) {
    const EnsClass& c = e->h[k];
    EnsClass* dc = e->d + k;
    const int u = e->vec_u[k];
    E_LAUNCH_VEC(k_e_smooth_dots, u, dim3(ENS_BANDS, c.B), dim3(256), dc);
    E_LAUNCH_VEC(k_e_dir, u, dim3(ENS_BANDS, c.B), dim3(256), dc);
    E_LAUNCH_VEC(k_e_update, u, dim3(ENS_BANDS, c.B), dim3(256), dc);
    E_LAUNCH_VEC(k_e_true, u, dim3(ENS_BANDS, c.B), dim3(256), dc);
    E_LAUNCH(k_e_true_fin, dim3(1), dim3(256), dc);
    return 0;
}

int ens_pred(cfd_ens* e, int k, int tx, int ty) {
    const EnsClass& c = e->h[k];
    const dim3 grid((c.n + 2 + tx - 1) / tx, (c.n + 2 + ty - 1) / ty, c.B);
    E_LAUNCH(k_e_pred, grid, dim3(tx, ty), e->d + k);
    return 0;
}

// one iteration with the complete V-cycle: legs down to level vlc, the LDS
// coarse end (one work-group per instance), legs up, the FP64 vector kernels
int ens_iteration_vcycle(cfd_ens* e) {
    for (int k = 0; k < e->ncls; ++k) {
        const EnsClass& c = e->h[k];
        EnsClass* dc = e->d + k;
        const int Lc = e->vlc[k];
        for (int l = 0; l < Lc; ++l) {
            E_LAUNCH(k_e_smooth, sw_grid(e, k, l), sw_block(e, k, l), dc, l, 0);
            E_LAUNCH(k_e_restrict, sw_grid(e, k, l + 1), sw_block(e, k, l + 1), dc, l);
        }
        E_LAUNCH(k_lv_vcoarse, dim3(c.B), dim3(1024), e->vargs[k]);
        for (int l = Lc - 1; l >= 1; --l) {
            E_LAUNCH(k_e_post, sw_grid(e, k, l), sw_block(e, k, l), dc, l);
            E_LAUNCH(k_e_smooth, sw_grid(e, k, l), sw_block(e, k, l), dc, l, 0);
        }
        E_LAUNCH(k_e_post, sw_grid(e, k, 0), sw_block(e, k, 0), dc, 0);
        TRY(ens_fcg_tail(e, k));
    }
    return 0;
}

int ens_iteration(cfd_ens* e) {
    if (e->vcycle) return ens_iteration_vcycle(e);
    for (int k = 0; k < e->ncls; ++k) {
        const EnsClass& c = e->h[k];
        EnsClass* dc = e->d + k;
        for (int l = 0; l < c.K; ++l) {
            E_LAUNCH(k_e_smooth, sw_grid(e, k, l), sw_block(e, k, l), dc, l, 0);
            E_LAUNCH(k_e_restrict, sw_grid(e, k, l + 1), sw_block(e, k, l + 1), dc, l);
        }
        E_LAUNCH(k_e_cmode, dim3(ENS_BANDS, c.B), dim3(256), dc);
        E_LAUNCH(k_e_net_in, dim3(4, c.B), dim3(256), dc);
    }
    if (e->use_net) {
        cfd_unet* u = e->net;
        const hipStream_t s0 = u->stream;
        u->stream = e->stream;
        const int rc = unet_forward(u, true);
        u->stream = s0;
        TRY(rc);
    }
    for (int k = 0; k < e->ncls; ++k) {
        const EnsClass& c = e->h[k];
        EnsClass* dc = e->d + k;
        const int K = c.K;
        // k_e_net_out's whole work-group computes the coarse statistics before
        // any thread leaves: it keeps 256 x 1 work-groups
        E_LAUNCH(k_e_net_out, dim3((unsigned)(c.lev[K].n * ((c.lev[K].n + 255) / 256)), c.B),
                 dim3(256), dc);
        E_LAUNCH(k_e_smooth, sw_grid(e, k, K), sw_block(e, k, K), dc, K, 1);
        E_LAUNCH(k_e_smooth, sw_grid(e, k, K), sw_block(e, k, K), dc, K, 0);
        for (int l = K - 1; l >= 1; --l) {
            E_LAUNCH(k_e_post, sw_grid(e, k, l), sw_block(e, k, l), dc, l);
            E_LAUNCH(k_e_smooth, sw_grid(e, k, l), sw_block(e, k, l), dc, l, 0);
        }
        E_LAUNCH(k_e_post, sw_grid(e, k, 0), sw_block(e, k, 0), dc, 0);
        TRY(ens_fcg_tail(e, k));
    }
    return 0;
}

void ens_drop_graph(cfd_ens* e) {
    if (e->graph_x) hipGraphExecDestroy(e->graph_x);
    if (e->graph) hipGraphDestroy(e->graph);
    e->graph_x = nullptr;
    e->graph = nullptr;
    e->cap_graph = 0;
}

int ens_notdone(cfd_ens* e, int* total) {
    HIPC(hipMemcpyAsync(e->notdone_h, e->notdone, e->ncls * sizeof(int), hipMemcpyDeviceToHost,
                        e->stream));
    HIPC(hipStreamSynchronize(e->stream));
    int t = 0;
    for (int k = 0; k < e->ncls; ++k) t += e->notdone_h[k];
    *total = t;
    return 0;
}

// before the solve: boundary conditions, dt / gamma, predictor, right-hand
// side, initial iterate and residual of every class
int ens_pre(cfd_ens* e) {
    for (int k = 0; k < e->ncls; ++k) {
        const EnsClass& c = e->h[k];
        EnsClass* dc = e->d + k;
        E_LAUNCH(k_e_bc, dim3(c.B), dim3(256), dc);
        E_LAUNCH(k_e_maxuv, dim3(ENS_BANDS, c.B), dim3(256), dc);
        E_LAUNCH(k_e_dt, dim3(1), dim3(256), dc);
        TRY(ens_pred(e, k, e->pred_tx[k], e->pred_ty[k]));
        E_LAUNCH(k_e_rhs, dim3(ENS_BANDS, c.B), dim3(256), dc);
        if (e->x0_extrap) E_LAUNCH(k_e_x0, dim3(ENS_BANDS, c.B), dim3(256), dc);
        E_LAUNCH(k_e_res, dim3(ENS_BANDS, c.B), dim3(256), dc);
        E_LAUNCH(k_e_init, dim3(1), dim3(256), dc);
    }
    return 0;
}

// after the solve: pressure, correction, temperature, monitors, state
// advance of every class, then the step's end (k_e_cap decided whether the
// step is complete; k_e_flip moves to the next step's buffers)
int ens_post(cfd_ens* e) {
    E_LAUNCH(k_e_cap, dim3(1), dim3(256), e->d, e->ncls);
    for (int k = 0; k < e->ncls; ++k) {
        const EnsClass& c = e->h[k];
        EnsClass* dc = e->d + k;
        E_LAUNCH(k_e_xsum, dim3(ENS_BANDS, c.B), dim3(256), dc);
        E_LAUNCH(k_e_pwrite, dim3(ENS_BANDS, c.B), dim3(256), dc);
        E_LAUNCH(k_e_correct, dim3(ENS_BANDS, c.B), dim3(256), dc);
        E_LAUNCH(k_e_bc, dim3(c.B), dim3(256), dc);
        const dim3 gt((c.n + 2 + e->pred_tx[k] - 1) / e->pred_tx[k],
                      (c.n + 2 + e->pred_ty[k] - 1) / e->pred_ty[k], c.B);
        E_LAUNCH(k_e_temp, gt, dim3(e->pred_tx[k], e->pred_ty[k]), dc);
        E_LAUNCH(k_e_tghost, dim3(c.B), dim3(256), dc);
        E_LAUNCH(k_e_record, dim3(ENS_BANDS, c.B), dim3(256), dc);
        E_LAUNCH(k_e_advance, dim3(1), dim3(256), dc);
    }
    E_LAUNCH(k_e_flip, dim3(1), dim3(64), e->ctl);
    return 0;
}

// a step with direct launches: iterations until every instance has finished
// (the host polls after the predicted count, then after each iteration)
int ens_step_direct(cfd_ens* e) {
    TRY(ens_pre(e));
    int batch = std::max(1, e->prev_iters - 1), launched = 0, left = 1, maxit = 0;
    for (int k = 0; k < e->ncls; ++k) maxit = std::max(maxit, e->h[k].maxiter);
    for (;;) {
        for (int q = 0; q < batch; ++q) TRY(ens_iteration(e));
        launched += batch;
        TRY(ens_notdone(e, &left));
        if (left == 0 || launched > maxit + 2) break;
        batch = 1;
    }
    e->prev_iters = launched;
    return ens_post(e);
}

// the step graph with `cap` iterations (captured once per cap)
int ens_graph_ready(cfd_ens* e) {
    if (!e->graph_x || e->cap_graph != e->cap) {
        ens_drop_graph(e);
        HIPC(hipStreamBeginCapture(e->stream, hipStreamCaptureModeRelaxed));
        int rc = ens_pre(e);
        for (int q = 0; q < e->cap && !rc; ++q) rc = ens_iteration(e);
        if (!rc) rc = ens_post(e);
        hipGraph_t g = nullptr;
        const hipError_t err = hipStreamEndCapture(e->stream, &g);
        if (rc) {
            if (g) hipGraphDestroy(g);
            return rc;
        }
        HIPC(err);
        e->graph = g;
        HIPC(hipGraphInstantiate(&e->graph_x, g, nullptr, nullptr, 0));
        e->cap_graph = e->cap;
    }
    return 0;
}

int ens_step_graph(cfd_ens* e) {
    TRY(ens_graph_ready(e));
    HIPC(hipGraphLaunch(e->graph_x, e->stream));
    return 0;
}

int ens_read_ctl(cfd_ens* e, EnsCtl* h) {
    HIPC(hipMemcpyAsync(h, e->ctl, sizeof(EnsCtl), hipMemcpyDeviceToHost, e->stream));
    HIPC(hipStreamSynchronize(e->stream));
    return 0;
}

// completed steps on the device (every instance advances in lockstep)
int ens_device_steps(cfd_ens* e, long long* n) {
    EnsState s;
    HIPC(hipMemcpyAsync(&s, e->h[0].st, sizeof(s), hipMemcpyDeviceToHost, e->stream));
    HIPC(hipStreamSynchronize(e->stream));
    *n = s.nstep;
    return 0;
}

// `count` steps: one graph launch each; launches that only continued a
// stalled solve are made up until the device has completed them all; the
// cap then follows the iteration count of the last step (+1)
int ens_steps(cfd_ens* e, int count) {
    if (!e->use_graph) {
        for (int s = 0; s < count; ++s) TRY(ens_step_direct(e));
        HIPC(hipStreamSynchronize(e->stream));
        e->steps += count;
        return 0;
    }
    long long n0 = 0;
    TRY(ens_device_steps(e, &n0));
    const long long target = n0 + count;
    for (int s = 0; s < count; ++s) TRY(ens_step_graph(e));
    long long n = 0;
    int maxit = 0;
    for (int k = 0; k < e->ncls; ++k) maxit = std::max(maxit, e->h[k].maxiter);
    const long long limit = (long long)count * (maxit / std::max(1, e->cap) + 2);
    long long extra = 0;
    for (;;) {
        TRY(ens_device_steps(e, &n));
        if (n >= target) break;
        if (++extra > limit)
            return set_err(CFD_ERR_STATE, "ensemble steps did not complete (%lld of %d)", n - n0,
                           count);
        TRY(ens_step_graph(e));
    }
    EnsCtl h;
    TRY(ens_read_ctl(e, &h));
    // the last step's count + 1: the counts fall as the flow settles, and a
    // step that needs more only continues in one more launch (+2 idled one
    // more iteration per step: 12.2 against 11.9 ms at 256 cavities)
    const int want = std::min(maxit, std::max(1, h.last_max + 1));
    h.max_call = 0;
    HIPC(hipMemcpy(&e->ctl->max_call, &h.max_call, sizeof(int), hipMemcpyHostToDevice));
    e->cap = want;
    e->steps += count;
    return ens_graph_ready(e);   // the next call starts without a capture
}

// Autotuner, run at creation before any step, on the class's own arrays:
//   - the predictor's work-group shape per class (it writes only F, G and the
//     tendency buffer of this parity, all zero from zero fields);
//   - per class and level the work-group shape of the pointwise V-cycle
//     kernels (smoothing, restriction, prolongation), timed on the level's
//     sweep;
//   - per class the cells per thread and batch of the FP64 vector kernels
//     (the fine sweep with the dot products, direction, update, true
//     residual), timed on the first two.
// Each candidate is the fastest of 5 timed launches after one untimed; every
// candidate computes the same values.  The timed kernels write the levels'
// b, x, t, the directions, q, the partial sums and the solver state, which
// are cleared afterwards: the class starts from its state at creation.
int ens_autotune(cfd_ens* e) {
    static const int pcand[][2] = {{64, 4}, {128, 2}, {256, 1}, {64, 8}, {128, 4}, {32, 8}};
    static const int scand[][2] = {{256, 1}, {128, 2}, {64, 4}, {32, 8}, {16, 16}};
    static const int ucand[] = {8, 2, 4, 16};   // the default first
    hipEvent_t a, b;
    HIPC(hipEventCreate(&a));
    HIPC(hipEventCreate(&b));
    auto timed = [&](auto launch, float* ms) -> int {
        TRY(launch());   // warm
        HIPC(hipEventRecord(a, e->stream));
        for (int r = 0; r < 5; ++r) TRY(launch());
        HIPC(hipEventRecord(b, e->stream));
        HIPC(hipEventSynchronize(b));
        HIPC(hipEventElapsedTime(ms, a, b));
        return 0;
    };
    int rc = 0;
    for (int k = 0; k < e->ncls && !rc; ++k) {
        const EnsClass& c = e->h[k];
        EnsClass* dc = e->d + k;
        float best = 1e30f, ms = 0.f;
        for (auto& cd : pcand) {
            if ((rc = timed([&] { return ens_pred(e, k, cd[0], cd[1]); }, &ms))) break;
            if (ms < best) {
                best = ms;
                e->pred_tx[k] = cd[0];
                e->pred_ty[k] = cd[1];
            }
        }
        e->pred_us[k] = 1e3f * best / 5;
        for (int l = 0; l < c.nlev && !rc; ++l) {
            best = 1e30f;
            int btx = 256, bty = 1;
            for (auto& cd : scand) {
                e->sw_tx[k][l] = cd[0];
                e->sw_ty[k][l] = cd[1];
                if ((rc = timed([&] {
                         E_LAUNCH(k_e_smooth, sw_grid(e, k, l), sw_block(e, k, l), dc, l, 0);
                         return 0;
                     }, &ms)))
                    break;
                if (cd[1] == 1) e->sw_us0[k][l] = 1e3f * ms / 5;   // the default 256 x 1
                if (ms < best) {
                    best = ms;
                    btx = cd[0];
                    bty = cd[1];
                }
            }
            e->sw_tx[k][l] = btx;
            e->sw_ty[k][l] = bty;
            e->sw_us[k][l] = 1e3f * best / 5;
        }
        if (rc) break;
        best = 1e30f;
        int bu = 8;
        for (int u : ucand) {
            e->vec_u[k] = u;
            if ((rc = timed([&] {
                     E_LAUNCH_VEC(k_e_smooth_dots, e->vec_u[k], dim3(ENS_BANDS, c.B), dim3(256), dc);
                     E_LAUNCH_VEC(k_e_dir, e->vec_u[k], dim3(ENS_BANDS, c.B), dim3(256), dc);
                     return 0;
                 }, &ms)))
                break;
            if (u == 8) e->vec_us0[k] = 1e3f * ms / 5;
            if (ms < best) {
                best = ms;
                bu = u;
            }
        }
        e->vec_u[k] = bu;
        e->vec_us[k] = 1e3f * best / 5;
    }
    hipEventDestroy(a);
    hipEventDestroy(b);
    if (rc) return rc;
    // the class's state at creation
    for (int k = 0; k < e->ncls; ++k) {
        const EnsClass& c = e->h[k];
        const size_t N = (size_t)c.S * c.B;
        for (int l = 0; l < c.nlev; ++l) {
            const size_t nl = (size_t)c.lev[l].S * c.B * sizeof(float);
            HIPC(hipMemsetAsync(c.lev[l].b, 0, nl, e->stream));
            HIPC(hipMemsetAsync(c.lev[l].x, 0, nl, e->stream));
            HIPC(hipMemsetAsync(c.lev[l].t, 0, nl, e->stream));
        }
        for (double* p : {c.q, c.d0, c.d1}) HIPC(hipMemsetAsync(p, 0, N * sizeof(double), e->stream));
        HIPC(hipMemsetAsync(c.part, 0, (size_t)ENS_NSLOT * c.B * ENS_BANDS * ENS_NV * sizeof(double),
                            e->stream));
        HIPC(hipMemsetAsync(c.st, 0, (size_t)c.B * sizeof(EnsState), e->stream));
    }
    HIPC(hipStreamSynchronize(e->stream));
    return 0;
}

}  // namespace

extern "C" {

int cfd_ens_create(const cfd_ens_desc* d, cfd_ens** out) {
    if (!d || !out || d->nclass < 1 || d->nclass > CFD_ENS_MAX_CLASSES)
        return set_err(CFD_ERR_ARG, "ensemble needs 1..%d classes", CFD_ENS_MAX_CLASSES);
    cfd_ens* e = new cfd_ens();
    auto fail = [&](int rc) {
        for (void* p : e->allocs) hipFree(p);
        if (e->notdone_h) hipHostFree(e->notdone_h);
        if (e->stream) hipStreamDestroy(e->stream);
        delete e;
        return rc;
    };
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(set_err(CFD_ERR_HIP, "stream creation failed"));
    e->ncls = d->nclass;
    int rc = 0;
    int img = 0;
    for (int k = 0; k < e->ncls; ++k) {
        const cfd_ens_class_desc& cd = d->cls[k];
        EnsClass& c = e->h[k];
        std::memset(&c, 0, sizeof(c));
        if (cd.n < 32 || cd.n % 32 || (cd.n & (cd.n - 1)) || cd.B < 1 || cd.B > 4096)
            return fail(set_err(CFD_ERR_ARG, "class %d: n %d (a power of two >= 32), B %d", k,
                                cd.n, cd.B));
        if (cd.nlev < 2 || cd.nlev > ENS_MAXLEV || !cd.cinv)
            return fail(set_err(CFD_ERR_ARG, "class %d: hierarchy of %d levels", k, cd.nlev));
        c.n = cd.n;
        c.P = ((cd.n + 2 + 31) / 32) * 32;
        c.B = cd.B;
        c.S = (long)(cd.n + 2) * c.P;
        c.h = 1.0f / cd.n;
        c.hd = 1.0 / cd.n;
        c.ax = (float)cd.n * (float)cd.n;
        c.axd = (double)cd.n * cd.n;
        c.tau = d->tau;
        c.nlev = cd.nlev;
        int K = 0;
        while ((cd.n >> K) > 32) ++K;
        c.K = K;
        if (K >= cd.nlev) return fail(set_err(CFD_ERR_ARG, "class %d: no 32x32 level", k));
        c.omega = d->omega;
        c.wc = d->wc;
        c.rtol = d->rtol;
        c.maxiter = d->maxiter;
        c.one_a_one = cd.one_a_one;
        c.img0 = img;
        img += cd.B;
        double *nu, *lid, *kap;
        if ((rc = ealloc(e, &nu, cd.B)) || (rc = ealloc(e, &lid, cd.B)) || (rc = ealloc(e, &kap, cd.B)))
            return fail(rc);
        if (hipMemcpy(nu, cd.nu, cd.B * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(lid, cd.lid, cd.B * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(kap, cd.kappa, cd.B * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            return fail(set_err(CFD_ERR_HIP, "parameter upload failed"));
        c.nu = nu;
        c.lid = lid;
        c.kappa = kap;
        const size_t N = (size_t)c.S * cd.B;
        float** f32[] = {&c.u, &c.v, &c.F, &c.G, &c.hu[0], &c.hu[1], &c.hv[0], &c.hv[1],
                         &c.T[0], &c.T[1], &c.p};
        for (float** p : f32)
            if ((rc = ealloc(e, p, N))) return fail(rc);
        double** f64[] = {&c.x, &c.b, &c.r, &c.ro, &c.q, &c.d0, &c.d1};
        for (double** p : f64)
            if ((rc = ealloc(e, p, N))) return fail(rc);
        c.xp = nullptr;
        if (d->x0_extrapolate && (rc = ealloc(e, &c.xp, N))) return fail(rc);
        for (int l = 0; l < cd.nlev; ++l) {
            EnsLevel& L = c.lev[l];
            L.n = cd.n >> l;
            if (cd.lev_n[l] != L.n)
                return fail(set_err(CFD_ERR_ARG, "class %d level %d: %d cells per side, expected %d",
                                    k, l, cd.lev_n[l], L.n));
            L.P = l == 0 ? c.P : ((L.n + 2 + 31) / 32) * 32;
            L.S = (long)(L.n + 2) * L.P;
            float *b, *x, *t;
            if ((rc = ealloc(e, &b, (size_t)L.S * cd.B)) || (rc = ealloc(e, &x, (size_t)L.S * cd.B)) ||
                (rc = ealloc(e, &t, (size_t)L.S * cd.B)))
                return fail(rc);
            L.b = b;
            L.x = x;
            L.t = t;
            if (l > 0) {
                float* co[3];
                const float* src[3] = {cd.diag[l], cd.ce[l], cd.cn[l]};
                for (int q = 0; q < 3; ++q) {
                    if (!src[q]) return fail(set_err(CFD_ERR_ARG, "class %d level %d operator", k, l));
                    if ((rc = ealloc(e, &co[q], (size_t)L.S))) return fail(rc);
                    if (hipMemcpy2D(co[q] + L.P + 1, L.P * sizeof(float), src[q], L.n * sizeof(float),
                                    L.n * sizeof(float), L.n, hipMemcpyHostToDevice) != hipSuccess)
                        return fail(set_err(CFD_ERR_HIP, "operator upload failed"));
                }
                L.diag = co[0];
                L.cE = co[1];
                L.cN = co[2];
            }
        }
        const int nc = cd.lev_n[cd.nlev - 1] * cd.lev_n[cd.nlev - 1];
        c.ncoarse = nc;
        float* cinv;
        if ((rc = ealloc(e, &cinv, (size_t)nc * nc))) return fail(rc);
        if (hipMemcpy(cinv, cd.cinv, (size_t)nc * nc * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
            return fail(set_err(CFD_ERR_HIP, "coarse inverse upload failed"));
        c.cinv = cinv;
        EnsState* st;
        double* part;
        if ((rc = ealloc(e, &st, cd.B)) ||
            (rc = ealloc(e, &part, (size_t)ENS_NSLOT * cd.B * ENS_BANDS * ENS_NV)))
            return fail(rc);
        c.st = st;
        c.part = part;
    }
    e->nimg = img;
    if ((rc = ealloc(e, &e->img, (size_t)img * 1024)) || (rc = ealloc(e, &e->d, e->ncls)) ||
        (rc = ealloc(e, &e->notdone, e->ncls)) || (rc = ealloc(e, &e->ctl, 1)))
        return fail(rc);
    {
        EnsCtl h;
        std::memset(&h, 0, sizeof(h));
        h.x0ext = d->x0_extrapolate ? 1 : 0;
        if (hipMemcpy(e->ctl, &h, sizeof(h), hipMemcpyHostToDevice) != hipSuccess)
            return fail(set_err(CFD_ERR_HIP, "control upload failed"));
    }
    if (hipHostMalloc((void**)&e->notdone_h, CFD_ENS_MAX_CLASSES * sizeof(int), 0) != hipSuccess)
        return fail(set_err(CFD_ERR_HIP, "pinned allocation failed"));
    for (int k = 0; k < e->ncls; ++k) {
        e->h[k].img = e->img;
        e->h[k].notdone = e->notdone + k;
        e->h[k].ctl = e->ctl;
        e->pred_tx[k] = 256;
        e->pred_ty[k] = 1;
        for (int l = 0; l < ENS_MAXLEV; ++l) {
            e->sw_tx[k][l] = 256;
            e->sw_ty[k][l] = 1;
            e->sw_us[k][l] = e->sw_us0[k][l] = 0.f;
        }
        e->vec_u[k] = 8;
        e->vec_us[k] = e->vec_us0[k] = 0.f;
    }
    if (hipMemcpy(e->d, e->h, e->ncls * sizeof(EnsClass), hipMemcpyHostToDevice) != hipSuccess)
        return fail(set_err(CFD_ERR_HIP, "class upload failed"));
    if (d->autotune && (rc = ens_autotune(e))) return fail(rc);
    e->use_graph = d->use_graph ? 1 : 0;
    e->x0_extrap = d->x0_extrapolate ? 1 : 0;
    *out = e;
    return 0;
}

int cfd_ens_destroy(cfd_ens* e) {
    if (!e) return 0;
    hipStreamSynchronize(e->stream);
    ens_drop_graph(e);
    for (void* p : e->allocs) hipFree(p);
    if (e->notdone_h) hipHostFree(e->notdone_h);
    hipStreamDestroy(e->stream);
    delete e;
    return 0;
}

// the shared network: image batch nimg x 32 x 32 (stacked), or NULL for the
// constant-vector coarse correction alone
int cfd_ens_set_net(cfd_ens* e, cfd_unet* net, float sigma) {
    if (!e) return set_err(CFD_ERR_ARG, "null ensemble");
    if (net && (net->H != 32 * e->nimg || net->W != 32 || net->img_h != 32))
        return set_err(CFD_ERR_ARG, "network must be a batch of %d images of 32 x 32", e->nimg);
    HIPC(hipStreamSynchronize(e->stream));
    ens_drop_graph(e);
    e->net = net;
    e->use_net = net ? 1 : 0;
    for (int k = 0; k < e->ncls; ++k) {
        EnsClass& c = e->h[k];
        c.use_net = e->use_net;
        c.img = net ? net->xin : e->img;
        c.y = net ? net->y : nullptr;
        c.gain = sigma / (4.0f * c.ax);
    }
    HIPC(hipMemcpy(e->d, e->h, e->ncls * sizeof(EnsClass), hipMemcpyHostToDevice));
    return 0;
}

int cfd_ens_set_coarse(cfd_ens* e, int32_t mode) {
    if (!e) return set_err(CFD_ERR_ARG, "null ensemble");
    if (mode != 0 && mode != 1) return set_err(CFD_ERR_ARG, "coarse mode %d", mode);
    HIPC(hipStreamSynchronize(e->stream));
    ens_drop_graph(e);
    if (mode == 0) {
        e->vcycle = 0;
        return 0;
    }
    for (int k = 0; k < e->ncls; ++k) {
        const EnsClass& c = e->h[k];
        int lc = 1;
        while (lc < c.nlev - 1 && (long)c.lev[lc].n * c.lev[lc].n > 64L * 64) ++lc;
        long need = 0;
        for (int l = lc; l < c.nlev; ++l) need += (long)VC_ARR * (c.lev[l].n + 2) * (c.lev[l].n + 2);
        need += 8;   // the float4 alignment of the pool's two parts (vc_assign)
        if (c.nlev - lc > VC_MAXL || c.ncoarse > VC_MAXC || need > VC_POOL)
            return set_err(CFD_ERR_LIMIT, "class %d: the coarse levels do not fit the LDS coarse end", k);
        VcArgs& v = e->vargs[k];
        std::memset(&v, 0, sizeof(v));
        v.done = reinterpret_cast<const char*>(c.st) + offsetof(EnsState, done);
        v.done_stride = (long)sizeof(EnsState);
        v.cinv = c.cinv;
        v.omega = c.omega;
        v.wc = c.wc;
        v.periodic = 0;
        v.nl = c.nlev - lc;
        v.nc = c.ncoarse;
        for (int q = 0; q < v.nl; ++q) {
            const EnsLevel& G = c.lev[lc + q];
            VcLev& L = v.lev[q];
            L.nx = L.ny = G.n;
            L.pw = G.n + 2;
            L.sz = (G.n + 2) * L.pw;
            L.P = G.P;
            L.m = ((1ull << 32) + (unsigned long long)G.n - 1) / (unsigned long long)G.n;
            L.istride = G.S;
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
        TRY(ealloc(e, &dimg, img.size()));
        HIPC(hipMemcpy(dimg, img.data(), img.size() * sizeof(float), hipMemcpyHostToDevice));
        v.image = reinterpret_cast<const float4*>(dimg);
        e->vlc[k] = lc;
    }
    e->vcycle = 1;
    return 0;
}

int cfd_ens_step(cfd_ens* e, int32_t count) {
    if (!e || count < 0) return set_err(CFD_ERR_ARG, "bad arguments");
    return ens_steps(e, count);
}

// fields of one instance in the reference layout: 0 u ((n+1) x (n+2)),
// 1 v ((n+2) x (n+1)), 2 p ((n+2) x (n+2)), 3 T ((n+2) x (n+2)); float
int cfd_ens_download(cfd_ens* e, int32_t cls, int32_t inst, int32_t field, float* host) {
    if (!e || !host || cls < 0 || cls >= e->ncls || inst < 0 || inst >= e->h[cls].B)
        return set_err(CFD_ERR_ARG, "bad arguments");
    const EnsClass& c = e->h[cls];
    const float* src;
    int rows, cols;
    switch (field) {
    case 0: src = c.u; rows = c.n + 1; cols = c.n + 2; break;
    case 1: src = c.v; rows = c.n + 2; cols = c.n + 1; break;
    case 2: src = c.p; rows = c.n + 2; cols = c.n + 2; break;
    case 3: src = c.T[0]; rows = c.n + 2; cols = c.n + 2; break;
    default: return set_err(CFD_ERR_ARG, "unknown field %d", field);
    }
    HIPC(hipStreamSynchronize(e->stream));
    if (field == 3) {   // the current temperature buffer has the device's parity
        EnsCtl h;
        TRY(ens_read_ctl(e, &h));
        src = c.T[h.par];
    }
    HIPC(hipMemcpy2D(host, cols * sizeof(float), src + inst * c.S, c.P * sizeof(float),
                     cols * sizeof(float), rows, hipMemcpyDeviceToHost));
    return 0;
}

int cfd_ens_stats(cfd_ens* e, int32_t cls, int32_t inst, cfd_ens_stats_t* out) {
    if (!e || !out || cls < 0 || cls >= e->ncls || inst < 0 || inst >= e->h[cls].B)
        return set_err(CFD_ERR_ARG, "bad arguments");
    HIPC(hipStreamSynchronize(e->stream));
    EnsState s;
    HIPC(hipMemcpy(&s, e->h[cls].st + inst, sizeof(s), hipMemcpyDeviceToHost));
    out->t = s.t;
    out->dt = s.dt;
    out->gamma = s.gamma;
    out->kinetic_energy = s.ke;
    out->divergence = s.divmax;
    out->residual = s.residual;
    out->relres = s.relres;
    out->steps = s.nstep;
    out->iters = s.iters;
    out->info = s.info;
    out->pred_tx = e->pred_tx[cls];
    out->pred_ty = e->pred_ty[cls];
    out->pred_us = e->pred_us[cls];
    EnsCtl h;
    TRY(ens_read_ctl(e, &h));
    out->sweeps_total = h.iters;
    return 0;
}

int cfd_ens_tuning(cfd_ens* e, int32_t cls, cfd_ens_tuning_t* out) {
    if (!e || !out || cls < 0 || cls >= e->ncls) return set_err(CFD_ERR_ARG, "bad arguments");
    std::memset(out, 0, sizeof(*out));
    out->pred_tx = e->pred_tx[cls];
    out->pred_ty = e->pred_ty[cls];
    out->pred_us = e->pred_us[cls];
    out->nlev = e->h[cls].nlev;
    for (int l = 0; l < e->h[cls].nlev; ++l) {
        out->sweep_tx[l] = e->sw_tx[cls][l];
        out->sweep_ty[l] = e->sw_ty[cls][l];
        out->sweep_us[l] = e->sw_us[cls][l];
        out->sweep_default_us[l] = e->sw_us0[cls][l];
    }
    out->vec_batch = e->vec_u[cls];
    out->vec_us = e->vec_us[cls];
    out->vec_default_us = e->vec_us0[cls];
    return 0;
}

}  // extern "C"
