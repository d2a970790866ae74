"""The U-Net's matrix-core kernels on the GPU against the FP64 oracle
(oracle/unet_oracle.py): the bf16 MFMA operand layout, the forward pass, the
parameter gradient and the Adam update."""

import numpy as np
import pytest

from oracle import unet_oracle as U

pytestmark = pytest.mark.gpu


def _dev(a):
    from cfd2d_amd._lib import DeviceArray

    return DeviceArray(np.ascontiguousarray(a, dtype=np.float32))


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-30))


def test_mfma_operand_layout(gpu):
    from cfd2d_amd.unet import mfma_layout_check

    rng = np.random.default_rng(0)
    A = U.bf16(rng.standard_normal((16, 16)))
    B = U.bf16(rng.standard_normal((16, 16)))
    D = mfma_layout_check(A, B)
    assert np.allclose(D, A @ B, rtol=1e-5, atol=1e-5)
    # a transposed operand would not pass
    assert not np.allclose(D, A.T @ B, rtol=1e-2, atol=1e-2)


def test_mfma_double_k_pairing(gpu):
    """Two 16-deep K steps through one v_mfma_f32_16x16x32_bf16, each lane
    passing its values of the first step and then of the second in A and B
    alike (as the convolutions do): the product over all 32 K positions."""
    from cfd2d_amd.unet import mfma_layout_check

    rng = np.random.default_rng(1)
    A = U.bf16(rng.standard_normal((16, 32)))
    B = U.bf16(rng.standard_normal((32, 16)))
    D = mfma_layout_check(A, B)
    assert np.allclose(D, A @ B, rtol=1e-5, atol=1e-5)
    # either half alone, or the halves crossed, would not pass
    assert not np.allclose(D, A[:, :16] @ B[:16], rtol=1e-2, atol=1e-2)
    crossed = A[:, :16] @ B[16:] + A[:, 16:] @ B[:16]
    assert not np.allclose(D, crossed, rtol=1e-2, atol=1e-2)


def _params(seed):
    p = U.init_params(seed).astype(np.float64)
    rng = np.random.default_rng(seed + 100)
    p += 0.02 * rng.standard_normal(p.size)
    return p.astype(np.float32)


@pytest.mark.parametrize("H,W", [(32, 48), (64, 64), (40, 24)])
def test_forward_matches_oracle(gpu, H, W):
    from cfd2d_amd.unet import UNet

    p = _params(1)
    rng = np.random.default_rng(H * W)
    x = rng.standard_normal((H, W)).astype(np.float32)
    net = UNet(H, W, p)
    dx, dy = _dev(x), _dev(np.zeros((H, W)))
    net.forward(dx, dy)
    y = dy.get()
    y_bf, _ = U.forward(p, x, bf16_emulate=True)
    y_64, _ = U.forward(p, x, bf16_emulate=False)
    print(f"{H}x{W}: rel err vs bf16-emulating oracle {_rel(y, y_bf):.2e}, vs FP64 {_rel(y, y_64):.2e}")
    assert _rel(y, y_bf) < 1e-2
    assert _rel(y, y_64) < 3e-2
    net.close()


@pytest.mark.parametrize("H,W,masked", [(32, 48, False), (64, 64, True)])
def test_gradient_matches_oracle(gpu, H, W, masked):
    from cfd2d_amd.unet import UNet

    p = _params(2)
    rng = np.random.default_rng(7)
    x = rng.standard_normal((H, W)).astype(np.float32)
    t = rng.standard_normal((H, W)).astype(np.float32)
    m = (rng.uniform(size=(H, W)) > 0.3).astype(np.float32) if masked else None
    tn = float(np.sum((1.0 if m is None else m) * t.astype(np.float64) ** 2))
    net = UNet(H, W, p)
    net.zero_grad()
    loss = net.loss_grad(_dev(x), _dev(t), None if m is None else _dev(m), tn)
    g = net.get_grad().astype(np.float64)
    loss_ref, g_ref = U.loss_grad(p, x, t, m, tn, bf16_emulate=True)
    assert loss == pytest.approx(loss_ref, rel=2e-3)
    offs, _ = U.offsets()
    worst = 0.0
    for l, (wo, bo) in enumerate(offs):
        co = U.LAYERS[l][3]
        ew, eb = _rel(g[wo:bo], g_ref[wo:bo]), _rel(g[bo:bo + co], g_ref[bo:bo + co])
        worst = max(worst, ew, eb)
        assert ew < 3e-2 and eb < 3e-2, (l, ew, eb)
    print(f"{H}x{W}: loss {loss:.6f} vs {loss_ref:.6f}, worst per-layer gradient rel err {worst:.2e}")
    # the whole gradient against the pure FP64 oracle (no bf16 emulation)
    _l64, g64 = U.loss_grad(p, x, t, m, tn, bf16_emulate=False)
    assert _rel(g, g64) < 5e-2
    net.close()


def test_adam_update_matches_oracle(gpu):
    from cfd2d_amd.unet import UNet

    H, W = 32, 32
    p = _params(3)
    rng = np.random.default_rng(8)
    x, t = rng.standard_normal((2, H, W)).astype(np.float32)
    net = UNet(H, W, p)
    dx, dt = _dev(x), _dev(t)
    w = p.astype(np.float64)
    m1, m2 = np.zeros_like(w), np.zeros_like(w)
    for step in (1, 2, 3):
        net.zero_grad()
        net.loss_grad(dx, dt, None, 1.0)
        g = net.get_grad().astype(np.float64)
        net.adam(1e-3)
        U.adam(w, g, m1, m2, step, 1e-3)
        wg = net.get_params().astype(np.float64)
        assert np.abs(wg - w).max() < 1e-6, step
        assert np.abs(net.get_grad()).max() == 0.0     # adam clears the gradient


def test_training_reduces_the_loss(gpu):
    """Fit the network to a fixed smooth map (the inverse 5-point Laplacian of
    random right-hand sides on a 32 x 32 grid): the loss falls by 5x."""
    from cfd2d_amd.unet import UNet

    H = W = 32
    rng = np.random.default_rng(9)
    n = H * W
    Lap = (4 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1) - np.eye(n, k=W) - np.eye(n, k=-W))
    for i in range(1, H):
        Lap[i * W, i * W - 1] = Lap[i * W - 1, i * W] = 0.0
    Linv = np.linalg.inv(Lap)
    pairs = []
    for _ in range(8):
        r = rng.standard_normal(n)
        e = Linv @ r
        s = np.sqrt(np.mean(r ** 2))
        pairs.append((_dev((r / s).reshape(H, W)), _dev((e / s).reshape(H, W)),
                      float(np.sum((e / s) ** 2))))
    net = UNet(H, W, seed=0)
    first = None
    for it in range(150):
        tot = 0.0
        for x, t, tn in pairs:
            tot += net.loss_grad(x, t, None, tn)
        net.adam(2e-3, 1.0 / len(pairs))
        first = tot / len(pairs) if first is None else first
    last = tot / len(pairs)
    print(f"loss {first:.4f} -> {last:.4f}")
    assert last < first / 5


@pytest.mark.parametrize("nimg", [5, 70])
def test_batched_32x32_forward(gpu, nimg):
    """Stacked batches of 32 x 32 images (the learned solver's coarse level;
    70 images take the one-work-group-per-image kernel with activations in
    LDS): every image against the bf16-emulating oracle and against its own
    single-image pass (per-layer kernels)."""
    import ctypes

    from cfd2d_amd import _lib
    from cfd2d_amd.unet import UNet

    p = _params(4)
    rng = np.random.default_rng(11)
    xs = rng.standard_normal((nimg, 32, 32)).astype(np.float32)
    net = UNet(32 * nimg, 32, p)
    _lib.check(_lib.lib().cfd_unet_set_batch(net.handle, 32), "batch")
    dx, dy = _dev(xs.reshape(32 * nimg, 32)), _dev(np.zeros((32 * nimg, 32)))
    net.forward(dx, dy)
    y = dy.get().reshape(nimg, 32, 32)
    one = UNet(32, 32, p)
    for k in range(0, nimg, max(1, nimg // 6)):
        # bf16 activations: a different fp32 summation order of a convolution
        # flips some roundings, which 15 layers compound to ~1% on noise inputs
        y_bf, _ = U.forward(p, xs[k], bf16_emulate=True)
        assert _rel(y[k], y_bf) < 2e-2, k
        d1, d2 = _dev(xs[k]), _dev(np.zeros((32, 32)))
        one.forward(d1, d2)
        if nimg < 64:
            assert np.array_equal(d2.get(), y[k])
        else:   # other kernel, other summation order
            assert _rel(y[k], d2.get()) < 2e-2
    net.close()
    one.close()
