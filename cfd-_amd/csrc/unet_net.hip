// The U-Net preconditioner network (forward, backward, Adam) built from the
// kernels of unet.hip.  Four levels, 3x3 convolutions, channels 16-32-64-64,
// ReLU, 2x2 average pooling, nearest upsampling with skip concatenation:
//
//   in(1)->16 ->16 | pool ->32 ->32 | pool ->64 ->64 | pool ->64 ->64
//   up + skip: (64+64)->64 ->64 | (64+32)->32 ->32 | (32+16)->16 ->16 | ->1
//
// Parameters are FP32 masters (flat vector) with BF16 copies of the MFMA
// layers' weights ([tap][Cout][Cin]); activations are BF16 NHWC; gradients
// FP32.  Image size H x W with H, W multiples of 8.

namespace unet {

enum LType { L_IN = 0, L_MFMA = 1, L_OUT = 2 };

struct Layer {
    int type, cin1, cin2, cout, level;
    long woff, boff;   // offsets into the flat parameter vector
};

// input tensor ids: 0 = none; activations indexed by name below
enum Act {
    A0A, A0B, P1, A1A, A1B, P2, A2A, A2B, P3, A3A, A3B, U2, B2A, B2B, U1, B1A, B1B, U0, B0A,
    B0B, NACT
};
static const int ACT_C[NACT] = {16, 16, 16, 32, 32, 32, 64, 64, 64, 64, 64, 64, 64, 64, 64,
                                32, 32, 32, 16, 16};
static const int ACT_L[NACT] = {0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 3, 2, 2, 2, 1, 1, 1, 0, 0, 0};

struct LayerIO {
    int x1, x2, y;   // activation ids (x1 = -1: the network input)
};

static const int NL = 15;
static const int LDEF[NL][5] = {
    // type, cin1, cin2, cout, level
    {L_IN, 1, 0, 16, 0},   {L_MFMA, 16, 0, 16, 0}, {L_MFMA, 16, 0, 32, 1},
    {L_MFMA, 32, 0, 32, 1}, {L_MFMA, 32, 0, 64, 2}, {L_MFMA, 64, 0, 64, 2},
    {L_MFMA, 64, 0, 64, 3}, {L_MFMA, 64, 0, 64, 3}, {L_MFMA, 64, 64, 64, 2},
    {L_MFMA, 64, 0, 64, 2}, {L_MFMA, 64, 32, 32, 1}, {L_MFMA, 32, 0, 32, 1},
    {L_MFMA, 32, 16, 16, 0}, {L_MFMA, 16, 0, 16, 0}, {L_OUT, 16, 0, 1, 0}};
static const LayerIO LIO[NL] = {
    {-1, -1, A0A}, {A0A, -1, A0B}, {P1, -1, A1A}, {A1A, -1, A1B}, {P2, -1, A2A},
    {A2A, -1, A2B}, {P3, -1, A3A}, {A3A, -1, A3B}, {U2, A2B, B2A}, {B2A, -1, B2B},
    {U1, A1B, B1A}, {B1A, -1, B1B}, {U0, A0B, B0A}, {B0A, -1, B0B}, {B0B, -1, -1}};

// pixels per work-group in the weight-gradient partial sums
constexpr long WG_PPB = 256;    // matrix-core layers
constexpr long SC_PPB = 1024;   // scalar in/out layers and biases

inline long param_count() {
    long n = 0;
    for (int l = 0; l < NL; ++l) n += 9L * (LDEF[l][1] + LDEF[l][2]) * LDEF[l][3] + LDEF[l][3];
    return n;
}

}  // namespace unet

struct cfd_unet {
    int H, W;
    int hs[4], ws[4];
    long npar;
    unet::Layer L[unet::NL];
    float* w = nullptr;             // fp32 masters
    unsigned short* wb = nullptr;   // bf16 copy
    unsigned short* wsw = nullptr;  // bf16 matrix-core weights in lane order (k_swizzle_b)
    unsigned short* wflip = nullptr;  // flipped/transposed bf16 weights (one layer at a time)
    float *g = nullptr, *m1 = nullptr, *m2 = nullptr;
    unsigned short* act[unet::NACT] = {};
    float* dact[unet::NACT] = {};
    unsigned short* gbuf = nullptr;  // masked gradient (bf16), largest activation
    float* xin = nullptr;            // network input (fp32 H*W)
    float* y = nullptr;              // network output (fp32 H*W)
    float* dy = nullptr;
    float* part = nullptr;
    long part_n = 0;
    int steps = 0;
    int img_h = 0;                   // rows per image of a stacked batch (level 0)
    int loss_kind = 0;               // 0 masked squared error, 1 energy norm
    int loss_nx = 0, loss_ny = 0;
    float loss_eps = 0.f;
    hipStream_t stream = nullptr;
    const int* skip = nullptr;       // forward kernels return at once while *skip != 0
    std::vector<void*> allocs;
};

namespace {

template <typename T>
int ualloc(cfd_unet* u, T** p, size_t n) {
    void* q = nullptr;
    HIPC(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
    HIPC(hipMemset(q, 0, std::max<size_t>(n, 1) * sizeof(T)));
    HIPC(hipDeviceSynchronize());   // the fill is done before any stream uses the array
    u->allocs.push_back(q);
    *p = (T*)q;
    return 0;
}

inline unsigned nblk(long n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

// the lane-ordered copy of the matrix-core layers' bf16 weights (after every
// change of wb)
int unet_swizzle(cfd_unet* u) {
    for (int l = 0; l < unet::NL; ++l) {
        const unet::Layer& Ly = u->L[l];
        if (Ly.type != unet::L_MFMA) continue;
        const int cin = Ly.cin1 + Ly.cin2;
        hipLaunchKernelGGL(k_swizzle_b, dim3(nblk(9L * cin * Ly.cout)), dim3(256), 0, u->stream,
                           u->wb + Ly.woff, cin, Ly.cout, u->wsw + Ly.woff);
        HIPC(hipGetLastError());
    }
    return 0;
}

int conv_mfma(cfd_unet* u, const unsigned short* x1, int c1, const unsigned short* x2, int c2,
              const unsigned short* wt, const float* bias, int H, int W, int cout,
              unsigned short* yb, float* yf, float* yf2, int csplit, int mode,
              const int* skip, int rs1 = 0, int Hi = 0) {
    if (cout % 16 || c1 % 16 || c2 % 16 || cout > 128)
        return set_err(CFD_ERR_ARG, "unsupported channel counts %d+%d -> %d", c1, c2, cout);
    // one work-group per (tile, output group); two 16-channel output chunks per
    // group unless that leaves fewer than 1024 work-groups
    const long tiles = (long)H * ((W + 15) / 16);
    const int G = (cout % 32 == 0 && tiles * (cout / 32) >= 1024) ? 2 : 1;
    const long tasks = tiles * (cout / (16 * G));
    const dim3 grid((unsigned)tasks);
    if (G == 2)
        hipLaunchKernelGGL(k_conv3x3_mfma<2>, grid, dim3(256), 0, u->stream, x1, c1, x2, c2, wt,
                           bias, H, W, cout, yb, yf, yf2, csplit, mode, skip, rs1, Hi > 0 ? Hi : H);
    else
        hipLaunchKernelGGL(k_conv3x3_mfma<1>, grid, dim3(256), 0, u->stream, x1, c1, x2, c2, wt,
                           bias, H, W, cout, yb, yf, yf2, csplit, mode, skip, rs1, Hi > 0 ? Hi : H);
    HIPC(hipGetLastError());
    return 0;
}

const unsigned short* act_in(cfd_unet* u, int id) { return id < 0 ? nullptr : u->act[id]; }

// infer: the pooled and upsampled tensors are not stored; the convolutions
// that consume them read their sources through the same resampling (the
// results are identical; the backward pass needs the stored tensors)
int unet_forward(cfd_unet* u, bool infer = false) {
    using namespace unet;
    const int Hi0 = u->img_h > 0 ? u->img_h : u->H;
    if (infer && Hi0 == 32 && u->W == 32 && u->H / 32 >= 64) {
        // batches of >= 64 images of 32 x 32: the whole network in one
        // work-group per image (one CU each; for fewer images the per-layer
        // kernels, which spread every layer over many CUs, are faster)
        hipLaunchKernelGGL(k_unet32, dim3(u->H / 32), dim3(1024), 0, u->stream, u->xin, u->y,
                           u->wsw, u->w, u->skip);
        HIPC(hipGetLastError());
        return 0;
    }
    for (int l = 0; l < NL; ++l) {
        const Layer& Ly = u->L[l];
        const int H = u->hs[Ly.level], W = u->ws[Ly.level];
        const bool pool_in = l == 2 || l == 4 || l == 6;
        const bool up_in = l == 8 || l == 10 || l == 12;
        const int Hi = (u->img_h > 0 ? u->img_h : u->H) >> Ly.level;   // rows per image
        if (infer && Ly.type == L_MFMA && (pool_in || up_in)) {
            const int src = pool_in ? (l == 2 ? A0B : (l == 4 ? A1B : A2B))
                                    : (l == 8 ? A3B : (l == 10 ? B2B : B1B));
            TRY(conv_mfma(u, u->act[src], Ly.cin1, act_in(u, LIO[l].x2), Ly.cin2,
                          u->wb + Ly.woff, u->w + Ly.boff, H, W, Ly.cout, u->act[LIO[l].y],
                          nullptr, nullptr, Ly.cout, 3, u->skip, pool_in ? 1 : 2, Hi));
            continue;
        }
        // resampling feeding this layer
        if (l == 2 || l == 4 || l == 6) {
            const int src = l == 2 ? A0B : (l == 4 ? A1B : A2B);
            const int dst = l == 2 ? P1 : (l == 4 ? P2 : P3);
            const int lf = ACT_L[src];
            const long n = (long)H * W * ACT_C[src];
            hipLaunchKernelGGL(k_avgpool2, dim3(nblk(n)), dim3(256), 0, u->stream, u->act[src],
                               u->hs[lf], u->ws[lf], ACT_C[src], u->act[dst], u->skip);
        }
        if (l == 8 || l == 10 || l == 12) {
            const int src = l == 8 ? A3B : (l == 10 ? B2B : B1B);
            const int dst = l == 8 ? U2 : (l == 10 ? U1 : U0);
            const long n = (long)H * W * ACT_C[src];
            hipLaunchKernelGGL(k_upsample2, dim3(nblk(n)), dim3(256), 0, u->stream, u->act[src],
                               H, W, ACT_C[src], u->act[dst], u->skip);
        }
        if (Ly.type == L_IN) {
            const long n = (long)H * W * Ly.cout;
            hipLaunchKernelGGL(k_conv_in, dim3(nblk(n)), dim3(256), 0, u->stream, u->xin, H, W,
                               u->w + Ly.woff, u->w + Ly.boff, Ly.cout, u->act[LIO[l].y],
                               u->skip, Hi);
        } else if (Ly.type == L_OUT) {
            hipLaunchKernelGGL(k_conv_out, dim3(nblk((long)H * W)), dim3(256), 0, u->stream,
                               u->act[LIO[l].x1], H, W, Ly.cin1, u->w + Ly.woff, u->w + Ly.boff,
                               u->y, u->skip, Hi);
        } else {
            TRY(conv_mfma(u, u->act[LIO[l].x1], Ly.cin1, act_in(u, LIO[l].x2), Ly.cin2,
                          u->wb + Ly.woff, u->w + Ly.boff, H, W, Ly.cout, u->act[LIO[l].y],
                          nullptr, nullptr, Ly.cout, 3, u->skip, 0, Hi));
        }
        HIPC(hipGetLastError());
    }
    return 0;
}

// gradient partials of a layer (nb rows of `stride` floats, the first n
// used, starting at part + first) summed into the flat gradient at off
int sum_into(cfd_unet* u, int nb, long stride, long first, long n, long off) {
    hipLaunchKernelGGL(k_sum_partials, dim3(nblk(n)), dim3(256), 0, u->stream, u->part + first,
                       nb, stride, n, u->g + off);
    HIPC(hipGetLastError());
    return 0;
}

int unet_backward(cfd_unet* u) {
    using namespace unet;
    for (int a = 0; a < NACT; ++a) {
        const long n = (long)u->hs[ACT_L[a]] * u->ws[ACT_L[a]] * ACT_C[a];
        hipLaunchKernelGGL(k_fill_f32, dim3(nblk(n)), dim3(256), 0, u->stream, u->dact[a], n, 0.f);
    }
    for (int l = NL - 1; l >= 0; --l) {
        const Layer& Ly = u->L[l];
        const int H = u->hs[Ly.level], W = u->ws[Ly.level];
        const long npix = (long)H * W;
        const long ppb = unet::SC_PPB;
        const int nb = (int)((npix + ppb - 1) / ppb);
        if (Ly.type == L_OUT) {
            hipLaunchKernelGGL(k_scalar_layer_wgrad, dim3(nb), dim3(256), 0, u->stream, 1,
                               (const float*)nullptr, (const unsigned short*)nullptr, u->dy,
                               u->act[LIO[l].x1], H, W, Ly.cin1, ppb, u->part);
            HIPC(hipGetLastError());
            // each block's partial row: [9][C] weights, then C copies of sum(dy)
            TRY(sum_into(u, nb, 10L * Ly.cin1, 0, 9L * Ly.cin1, Ly.woff));
            TRY(sum_into(u, nb, 10L * Ly.cin1, 9L * Ly.cin1, 1, Ly.boff));
            hipLaunchKernelGGL(k_conv_out_bwd_data, dim3(nblk(npix * Ly.cin1)), dim3(256), 0,
                               u->stream, u->dy, H, W, Ly.cin1, u->w + Ly.woff,
                               u->dact[LIO[l].x1]);
            HIPC(hipGetLastError());
            continue;
        }
        // masked gradient G = d(out) * relu'(out)
        const int yid = LIO[l].y;
        const long ny = npix * Ly.cout;
        hipLaunchKernelGGL(k_relu_mask, dim3(nblk(ny)), dim3(256), 0, u->stream, u->dact[yid],
                           u->act[yid], ny, u->gbuf);
        HIPC(hipGetLastError());
        if (Ly.type == L_IN) {
            hipLaunchKernelGGL(k_scalar_layer_wgrad, dim3(nb), dim3(256), 0, u->stream, 0,
                               u->xin, u->gbuf, (const float*)nullptr,
                               (const unsigned short*)nullptr, H, W, Ly.cout, ppb, u->part);
            HIPC(hipGetLastError());
            // partial rows: [9][C] weights then [C] bias, contiguous per block
            // partial rows [9][C] weights then [C] bias = the layout at woff (boff = woff + 9C)
            TRY(sum_into(u, nb, 10L * Ly.cout, 0, 10L * Ly.cout, Ly.woff));
            continue;
        }
        const int cin = Ly.cin1 + Ly.cin2;
        // weight gradient: pixel chunks x (tap, output chunk) groups of 4 waves
        const int nbw = (int)((npix + unet::WG_PPB - 1) / unet::WG_PPB);
        const dim3 gw(nbw, (9 * (Ly.cout / 16) + 3) / 4);
        hipLaunchKernelGGL(k_conv3x3_wgrad, gw, dim3(256), 0, u->stream, u->gbuf, Ly.cout,
                           u->act[LIO[l].x1], Ly.cin1, act_in(u, LIO[l].x2), Ly.cin2, H, W,
                           (long)unet::WG_PPB, u->part);
        HIPC(hipGetLastError());
        TRY(sum_into(u, nbw, 9L * Ly.cout * cin, 0, 9L * Ly.cout * cin, Ly.woff));
        hipLaunchKernelGGL(k_bias_grad, dim3(nb), dim3(256), 0, u->stream, u->gbuf, npix,
                           Ly.cout, ppb, u->part);
        HIPC(hipGetLastError());
        TRY(sum_into(u, nb, Ly.cout, 0, Ly.cout, Ly.boff));
        // data gradient: conv of G with flipped, transposed weights (accumulated)
        const long nw = 9L * Ly.cout * cin;
        hipLaunchKernelGGL(k_flip_transpose, dim3(nblk(nw)), dim3(256), 0, u->stream,
                           u->wb + Ly.woff, Ly.cout, cin, u->wflip);
        HIPC(hipGetLastError());
        float* d1 = u->dact[LIO[l].x1];
        float* d2 = LIO[l].x2 >= 0 ? u->dact[LIO[l].x2] : nullptr;
        TRY(conv_mfma(u, u->gbuf, Ly.cout, nullptr, 0, u->wflip, nullptr, H, W, cin, nullptr, d1,
                      d2, Ly.cin1, 4, nullptr));
        // resampling that fed this layer, backwards
        if (l == 2 || l == 4 || l == 6) {
            const int src = l == 2 ? A0B : (l == 4 ? A1B : A2B);
            const int dst = l == 2 ? P1 : (l == 4 ? P2 : P3);
            const int lf = ACT_L[src];
            const long n = (long)u->hs[lf] * u->ws[lf] * ACT_C[src];
            hipLaunchKernelGGL(k_avgpool2_bwd, dim3(nblk(n)), dim3(256), 0, u->stream,
                               u->dact[dst], u->hs[lf], u->ws[lf], ACT_C[src], u->dact[src]);
        }
        if (l == 8 || l == 10 || l == 12) {
            const int src = l == 8 ? A3B : (l == 10 ? B2B : B1B);
            const int dst = l == 8 ? U2 : (l == 10 ? U1 : U0);
            const int lc = ACT_L[src];
            const long n = (long)u->hs[lc] * u->ws[lc] * ACT_C[src];
            hipLaunchKernelGGL(k_upsample2_bwd, dim3(nblk(n)), dim3(256), 0, u->stream,
                               u->dact[dst], H, W, ACT_C[src], u->dact[src]);
        }
        HIPC(hipGetLastError());
    }
    return 0;
}

}  // namespace

extern "C" {

int64_t cfd_unet_param_count(void) { return unet::param_count(); }

int cfd_unet_create(int32_t H, int32_t W, cfd_unet** out) {
    if (!out || H < 8 || W < 8 || H % 8 || W % 8)
        return set_err(CFD_ERR_ARG, "U-Net image size must be multiples of 8 (got %d x %d)", H, W);
    cfd_unet* u = new cfd_unet();
    u->H = H;
    u->W = W;
    for (int k = 0; k < 4; ++k) {
        u->hs[k] = H >> k;
        u->ws[k] = W >> k;
    }
    long off = 0;
    for (int l = 0; l < unet::NL; ++l) {
        unet::Layer& Ly = u->L[l];
        Ly.type = unet::LDEF[l][0];
        Ly.cin1 = unet::LDEF[l][1];
        Ly.cin2 = unet::LDEF[l][2];
        Ly.cout = unet::LDEF[l][3];
        Ly.level = unet::LDEF[l][4];
        Ly.woff = off;
        off += 9L * (Ly.cin1 + Ly.cin2) * Ly.cout;
        Ly.boff = off;
        off += Ly.cout;
        if (Ly.woff != img_layout::woff(l) || Ly.boff != img_layout::boff(l)) {
            delete u;
            return set_err(CFD_ERR_STATE, "layer %d offsets disagree with the one-image network", l);
        }
    }
    u->npar = off;
    auto fail = [&](int rc) {
        for (void* p : u->allocs) hipFree(p);
        delete u;
        return rc;
    };
    int rc = 0;
    const long n0 = (long)H * W;
    long maxact = 0;
    for (int a = 0; a < unet::NACT; ++a) {
        const long n = (long)u->hs[unet::ACT_L[a]] * u->ws[unet::ACT_L[a]] * unet::ACT_C[a];
        maxact = std::max(maxact, n);
        if ((rc = ualloc(u, &u->act[a], n)) || (rc = ualloc(u, &u->dact[a], n))) return fail(rc);
    }
    // partial-sum buffer: the largest of the per-layer needs
    long pn = (n0 + 255) / 256 + 1024;
    for (int l = 0; l < unet::NL; ++l) {
        const unet::Layer& Ly = u->L[l];
        const long np = (long)u->hs[Ly.level] * u->ws[Ly.level];
        const long cw = (np + unet::WG_PPB - 1) / unet::WG_PPB;
        const long cs = (np + unet::SC_PPB - 1) / unet::SC_PPB;
        pn = std::max(pn, cw * 9L * Ly.cout * (Ly.cin1 + Ly.cin2));
        pn = std::max(pn, cs * 10L * std::max(Ly.cout, Ly.cin1));
    }
    u->part_n = pn;
    if ((rc = ualloc(u, &u->w, u->npar)) || (rc = ualloc(u, &u->wb, u->npar)) ||
        (rc = ualloc(u, &u->wsw, u->npar)) ||
        (rc = ualloc(u, &u->wflip, 9L * 128 * 64)) || (rc = ualloc(u, &u->g, u->npar)) ||
        (rc = ualloc(u, &u->m1, u->npar)) || (rc = ualloc(u, &u->m2, u->npar)) ||
        (rc = ualloc(u, &u->gbuf, maxact)) || (rc = ualloc(u, &u->xin, n0)) ||
        (rc = ualloc(u, &u->y, n0)) || (rc = ualloc(u, &u->dy, n0)) ||
        (rc = ualloc(u, &u->part, u->part_n)))
        return fail(rc);
    *out = u;
    return 0;
}

int cfd_unet_destroy(cfd_unet* u) {
    if (!u) return 0;
    if (u->stream) hipStreamSynchronize(u->stream);
    hipDeviceSynchronize();
    for (void* p : u->allocs) hipFree(p);
    delete u;
    return 0;
}

int cfd_unet_set_params(cfd_unet* u, const float* host) {
    if (!u || !host) return set_err(CFD_ERR_ARG, "null argument");
    HIPC(hipMemcpy(u->w, host, u->npar * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_to_bf16, dim3(nblk(u->npar)), dim3(256), 0, u->stream, u->w, u->npar,
                       u->wb);
    HIPC(hipGetLastError());
    TRY(unet_swizzle(u));
    HIPC(hipMemset(u->m1, 0, u->npar * sizeof(float)));
    HIPC(hipMemset(u->m2, 0, u->npar * sizeof(float)));
    u->steps = 0;
    HIPC(hipDeviceSynchronize());
    return 0;
}

int cfd_unet_get_params(cfd_unet* u, float* host) {
    if (!u || !host) return set_err(CFD_ERR_ARG, "null argument");
    HIPC(hipDeviceSynchronize());
    HIPC(hipMemcpy(host, u->w, u->npar * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

int cfd_unet_get_grad(cfd_unet* u, float* host) {
    if (!u || !host) return set_err(CFD_ERR_ARG, "null argument");
    HIPC(hipDeviceSynchronize());
    HIPC(hipMemcpy(host, u->g, u->npar * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

int cfd_unet_zero_grad(cfd_unet* u) {
    if (!u) return set_err(CFD_ERR_ARG, "null argument");
    hipLaunchKernelGGL(k_fill_f32, dim3(nblk(u->npar)), dim3(256), 0, u->stream, u->g, u->npar, 0.f);
    HIPC(hipGetLastError());
    return 0;
}

// y = net(x) for device images x, y (fp32, H*W); x may be the network's own buffer
int cfd_unet_forward(cfd_unet* u, const float* x_dev, float* y_dev) {
    if (!u || !x_dev || !y_dev) return set_err(CFD_ERR_ARG, "null argument");
    const long n = (long)u->H * u->W;
    if (x_dev != u->xin) HIPC(hipMemcpyAsync(u->xin, x_dev, n * sizeof(float), hipMemcpyDeviceToDevice, u->stream));
    TRY(unet_forward(u, true));
    if (y_dev != u->y) HIPC(hipMemcpyAsync(y_dev, u->y, n * sizeof(float), hipMemcpyDeviceToDevice, u->stream));
    return 0;
}

// loss = sum mask (net(x) - t)^2 / sum mask t^2; gradients accumulate into the
// network's gradient vector.  x, t, mask: device fp32 images (mask may be NULL).
int cfd_unet_loss_grad(cfd_unet* u, const float* x_dev, const float* t_dev, const float* mask_dev,
                       double tnorm2, double* loss) {
    if (!u || !x_dev || !t_dev || !loss || !(tnorm2 > 0.0)) return set_err(CFD_ERR_ARG, "bad arguments");
    const long n = (long)u->H * u->W;
    HIPC(hipMemcpyAsync(u->xin, x_dev, n * sizeof(float), hipMemcpyDeviceToDevice, u->stream));
    TRY(unet_forward(u));
    const unsigned nb = nblk(n);
    if (u->loss_kind == 1)
        hipLaunchKernelGGL(k_energy, dim3(nb), dim3(256), 0, u->stream, u->y, t_dev, u->H, u->W,
                           u->loss_nx, u->loss_ny, u->loss_eps, (float)(1.0 / tnorm2), u->dy,
                           u->part);
    else
        hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, u->stream, u->y, t_dev, mask_dev, n,
                           (float)(1.0 / tnorm2), u->dy, u->part);
    HIPC(hipGetLastError());
    std::vector<float> lp(nb);
    HIPC(hipMemcpyAsync(lp.data(), u->part, nb * sizeof(float), hipMemcpyDeviceToHost, u->stream));
    HIPC(hipStreamSynchronize(u->stream));
    double s = 0.0;
    for (float v : lp) s += v;
    *loss = s;
    TRY(unet_backward(u));
    return 0;
}

int cfd_unet_set_batch(cfd_unet* u, int32_t image_rows) {
    if (!u || image_rows < 8 || image_rows % 8 || u->H % image_rows)
        return set_err(CFD_ERR_ARG, "batch image rows %d do not divide %d in multiples of 8",
                       image_rows, u ? u->H : 0);
    u->img_h = image_rows;
    return 0;
}

int cfd_unet_set_loss(cfd_unet* u, int32_t kind, int32_t nx, int32_t ny, float eps) {
    if (!u || kind < 0 || kind > 1) return set_err(CFD_ERR_ARG, "bad loss kind");
    if (kind == 1 && (nx < 1 || ny < 1 || nx > u->H || ny > u->W || !(eps >= 0.f)))
        return set_err(CFD_ERR_ARG, "energy loss region %d x %d in a %d x %d image", nx, ny, u->H,
                       u->W);
    u->loss_kind = kind;
    u->loss_nx = nx;
    u->loss_ny = ny;
    u->loss_eps = eps;
    return 0;
}

int cfd_unet_adam(cfd_unet* u, float lr, float scale) {
    if (!u) return set_err(CFD_ERR_ARG, "null argument");
    u->steps += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)u->steps), c2 = 1.f - powf(b2, (float)u->steps);
    if (scale != 1.0f) {
        // average over a batch: scale the accumulated gradient
        hipLaunchKernelGGL(k_scale_f32, dim3(nblk(u->npar)), dim3(256), 0, u->stream, u->g,
                           u->npar, scale);
    }
    hipLaunchKernelGGL(k_adam, dim3(nblk(u->npar)), dim3(256), 0, u->stream, u->w, u->g, u->m1,
                       u->m2, u->npar, lr, b1, b2, 1e-8f, c1, c2, u->wb);
    HIPC(hipGetLastError());
    TRY(unet_swizzle(u));
    TRY(cfd_unet_zero_grad(u));
    return 0;
}

int cfd_mfma_layout_check(const float* A, const float* B, float* D) {
    if (!A || !B || !D) return set_err(CFD_ERR_ARG, "null argument");
    float *a, *b, *d;
    HIPC(hipMalloc(&a, 256 * sizeof(float)));
    HIPC(hipMalloc(&b, 256 * sizeof(float)));
    HIPC(hipMalloc(&d, 256 * sizeof(float)));
    HIPC(hipMemcpy(a, A, 256 * sizeof(float), hipMemcpyHostToDevice));
    HIPC(hipMemcpy(b, B, 256 * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_mfma_layout_check, dim3(1), dim3(64), 0, 0, a, b, d);
    HIPC(hipGetLastError());
    HIPC(hipMemcpy(D, d, 256 * sizeof(float), hipMemcpyDeviceToHost));
    hipFree(a);
    hipFree(b);
    hipFree(d);
    return 0;
}

int cfd_mfma_layout_check_k32(const float* A, const float* B, float* D) {
    if (!A || !B || !D) return set_err(CFD_ERR_ARG, "null argument");
    float *a, *b, *d;
    HIPC(hipMalloc(&a, 512 * sizeof(float)));
    HIPC(hipMalloc(&b, 512 * sizeof(float)));
    HIPC(hipMalloc(&d, 256 * sizeof(float)));
    HIPC(hipMemcpy(a, A, 512 * sizeof(float), hipMemcpyHostToDevice));
    HIPC(hipMemcpy(b, B, 512 * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_mfma_layout_check_k32, dim3(1), dim3(64), 0, 0, a, b, d);
    HIPC(hipGetLastError());
    HIPC(hipMemcpy(D, d, 256 * sizeof(float), hipMemcpyDeviceToHost));
    hipFree(a);
    hipFree(b);
    hipFree(d);
    return 0;
}

}  // extern "C"
