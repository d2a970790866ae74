"""Pins of the CPU oracle of the U-Net (oracle/unet_oracle.py): brute-force
convolution, adjoint identities of pooling and upsampling, bfloat16 rounding,
and the backward pass against central finite differences."""

import numpy as np
import pytest

from oracle import unet_oracle as U


def test_conv_matches_brute_force():
    rng = np.random.default_rng(3)
    H, W, ci, co = 5, 7, 3, 4
    x = rng.standard_normal((H, W, ci))
    Wt = rng.standard_normal((9, co, ci))
    b = rng.standard_normal(co)
    y = U.conv3x3(x, Wt, b)
    ref = np.zeros((H, W, co))
    for h in range(H):
        for w in range(W):
            for c in range(co):
                s = b[c]
                for dh in (-1, 0, 1):
                    for dw in (-1, 0, 1):
                        hh, ww = h + dh, w + dw
                        if 0 <= hh < H and 0 <= ww < W:
                            s += Wt[3 * (dh + 1) + dw + 1, c] @ x[hh, ww]
                ref[h, w, c] = s
    assert np.allclose(y, ref, rtol=1e-13, atol=1e-13)


def test_conv_backward_is_the_adjoint():
    rng = np.random.default_rng(4)
    x = rng.standard_normal((6, 8, 3))
    Wt = rng.standard_normal((9, 5, 3))
    g = rng.standard_normal((6, 8, 5))
    dx, dW, db = U.conv3x3_bwd(x, Wt, g)
    # <conv(x), g> is linear in x and in W: both gradients reproduce it
    lin = np.sum(U.conv3x3(x, Wt, None) * g)
    assert np.isclose(np.sum(dx * x), lin, rtol=1e-12)
    assert np.isclose(np.sum(dW * Wt), lin, rtol=1e-12)
    assert np.allclose(db, g.sum(axis=(0, 1)))


def test_pool_and_upsample_adjoints():
    rng = np.random.default_rng(5)
    x = rng.standard_normal((8, 6, 2))
    c = rng.standard_normal((4, 3, 2))
    assert np.isclose(np.sum(U.pool2(x) * c), np.sum(x * U.pool2_bwd(c)), rtol=1e-13)
    assert np.isclose(np.sum(U.up2(c) * x), np.sum(c * U.up2_bwd(x)), rtol=1e-13)
    assert np.allclose(U.pool2(U.up2(c)), c)


def test_bf16_rounding_nearest_even():
    # 1 + 2^-8 is halfway between two bf16 numbers: ties go to the even one (1.0);
    # 1 + 3*2^-8 goes up to 1 + 2^-6
    v = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -2.5, 3.0e38], dtype=np.float32)
    r = U.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1 + 2 ** -6 and r[3] == -2.5
    assert np.all(np.abs(r - v) <= np.abs(v) * 2 ** -8)


def test_layout_and_init_match_the_host_package():
    from cfd2d_amd import unet

    assert U.param_count() == unet.param_count() == 302689
    assert [tuple(l[:4]) for l in U.LAYERS] == [tuple(l) for l in unet.LAYERS]
    assert np.array_equal(U.init_params(7), unet.init_params(7))


def test_backward_matches_finite_differences():
    p = U.init_params(1).astype(np.float64)
    rng = np.random.default_rng(0)
    p += 0.01 * rng.standard_normal(p.size)          # non-zero biases too
    x = rng.standard_normal((16, 24))
    t = rng.standard_normal((16, 24))
    mask = (rng.uniform(size=(16, 24)) > 0.2).astype(float)
    loss, g = U.loss_grad(p, x, t, mask)
    offs, _ = U.offsets()
    eps = 1e-6
    checked = 0
    for wo, bo in offs:
        for k in (wo, (wo + bo) // 2, bo - 1, bo):
            pp, pm = p.copy(), p.copy()
            pp[k] += eps
            pm[k] -= eps
            fd = (U.loss_grad(pp, x, t, mask)[0] - U.loss_grad(pm, x, t, mask)[0]) / (2 * eps)
            assert abs(fd - g[k]) <= 1e-5 * (abs(fd) + abs(g[k])) + 1e-10, (k, fd, g[k])
            checked += 1
    assert checked == 4 * len(U.LAYERS)
    assert loss > 0


def test_adam_first_step_is_sign_times_lr():
    w = np.array([1.0, -2.0, 3.0])
    g = np.array([0.5, -0.1, 0.0])
    m1, m2 = np.zeros(3), np.zeros(3)
    U.adam(w, g, m1, m2, 1, 0.01)
    # after one step the bias-corrected update is lr * g / (|g| + eps)
    assert np.allclose(w, [0.99, -1.99, 3.0], atol=1e-7)
