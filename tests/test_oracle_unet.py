"""Pins of the FP64 U-Net oracle: convolution against scipy's correlate,
pooling/upsampling against explicit index formulas, the convolution backward
by its adjoint identities, the operator against the textbook 7-point
Laplacian, and the whole network's parameter gradient against central finite
differences of its loss."""
import numpy as np
import scipy.ndimage as ndi
import scipy.sparse as sp

from oracle import unet_oracle as uo
from tests.unet_cases import synthetic_system


def test_bf16_round_known_values():
    v = np.array([1.0, 1 + 2 ** -7, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -2.5, 0.0, 3 * 2 ** -9 + 1], np.float32)
    want = np.array([1.0, 1 + 2 ** -7, 1.0, 1 + 2 ** -6, -2.5, 0.0, 1 + 2 ** -7])
    assert np.array_equal(uo.bf16_round(v), want)


def test_conv3_matches_scipy_correlate():
    rng = np.random.default_rng(3)
    n, cin, cout = 6, 3, 2
    x = rng.normal(size=(n, n, n, cin))
    W = rng.normal(size=(27, cout, cin))
    b = rng.normal(size=cout)
    y = uo.conv3(x, W, b)
    for co in range(cout):
        want = np.full((n, n, n), b[co])
        for ci in range(cin):
            k = W[:, co, ci].reshape(3, 3, 3)  # [dz][dy][dx]
            want += ndi.correlate(x[..., ci], k, mode="constant", cval=0.0)
        assert np.allclose(y[..., co], want, rtol=1e-12, atol=1e-12)


def test_conv3_backward_adjoint_identities():
    rng = np.random.default_rng(4)
    n, cin, cout = 5, 4, 3
    x = rng.normal(size=(n, n, n, cin))
    W = rng.normal(size=(27, cout, cin))
    g = rng.normal(size=(n, n, n, cout))
    dx, dW, db = uo.conv3_backward(x, W, g)
    y0 = uo.conv3(x, W, np.zeros(cout))
    # <conv(x), g> is bilinear: = <x, dx> = <W, dW>
    assert np.isclose((y0 * g).sum(), (x * dx).sum(), rtol=1e-12)
    assert np.isclose((y0 * g).sum(), (W * dW).sum(), rtol=1e-12)
    assert np.allclose(db, g.reshape(-1, cout).sum(axis=0))


def test_pool_and_up_index_formulas():
    rng = np.random.default_rng(5)
    x = rng.normal(size=(4, 4, 4, 2))
    p = uo.pool(x)
    for z in range(2):
        for y in range(2):
            for xx in range(2):
                blk = x[2 * z:2 * z + 2, 2 * y:2 * y + 2, 2 * xx:2 * xx + 2]
                assert np.allclose(p[z, y, xx], blk.reshape(-1, 2).mean(axis=0))
    u = uo.up(p)
    for z in range(4):
        for y in range(4):
            for xx in range(4):
                assert np.array_equal(u[z, y, xx], p[z // 2, y // 2, xx // 2])
    g = rng.normal(size=(4, 4, 4, 2))
    # adjoints: <up(p), g> = <p, up^T g>;  <pool(x), h> = <x, pool^T h>
    assert np.isclose((u * g).sum(), (p * uo.up_backward(g)).sum())
    h = rng.normal(size=p.shape)
    assert np.isclose((p * h).sum(), (x * uo.pool_backward(h)).sum())


def test_apply_A_is_7_point_laplacian():
    n = 5
    coef = np.zeros((n, n, n), np.uint16)
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    m = ((x > 0) * 1 | (x < n - 1) * 2 | (y > 0) * 4 | (y < n - 1) * 8 | (z > 0) * 16 | (z < n - 1) * 32)
    coef[...] = (6 | (m << 3)).astype(np.uint16)
    T = sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1])
    I = sp.identity(n)
    L = sp.kron(sp.kron(T, I), I) + sp.kron(sp.kron(I, T), I) + sp.kron(sp.kron(I, I), T)
    v = np.random.default_rng(6).normal(size=(n, n, n))
    assert np.allclose(uo.apply_A(coef, v).reshape(-1), L @ v.reshape(-1), rtol=1e-12, atol=1e-12)


def test_synthetic_system_is_symmetric():
    b, coef = synthetic_system(8, 1)
    rng = np.random.default_rng(7)
    u, v = rng.normal(size=(2, 8, 8, 8))
    assert np.isclose((uo.apply_A(coef, u) * v).sum(), (u * uo.apply_A(coef, v)).sum(), rtol=1e-12)


def test_unet_gradient_matches_finite_differences():
    n = 8
    b, coef = synthetic_system(n, 2)
    p = uo.init_like_library(11).astype(np.float64)
    layout, hw, hb, total = uo.param_layout()
    # small random biases so the ReLUs are not all symmetric around 0
    rng = np.random.default_rng(12)
    for w, bo, cin, cout, _ in layout:
        p[bo:bo + cout] = rng.normal(0.0, 0.05, cout)
    p[hb] = 0.01
    net = uo.Net(n, p, bf16=False)
    L, g = net.loss_grad(b, coef)
    assert L > 0
    idx = [hw + 3, hb]
    for w, bo, cin, cout, _ in layout:
        idx += list(rng.integers(w, bo, 3)) + [bo + int(rng.integers(0, cout))]
    # the loss is piecewise smooth (ReLU kinks): a step that straddles a kink
    # spoils the difference quotient, so the closer of two step sizes counts
    for i in idx:
        errs = []
        for eps in (1e-6, 1e-7):
            pp = p.copy()
            pp[i] += eps
            lp = uo.Net(n, pp).loss(b, coef)
            pp[i] -= 2 * eps
            lm = uo.Net(n, pp).loss(b, coef)
            fd = (lp - lm) / (2 * eps)
            errs.append(abs(fd - g[i]) / (max(1.0, abs(fd)) * 1e-6 + 1e-8))
        assert min(errs) <= 1.0, (i, errs, g[i])


def test_activations_vanish_outside_the_level_masks():
    # M0 = fluid; M1, M2 = any over 2x2x2 children (checked by brute force);
    # every stored activation is zero outside its level's mask, so changing
    # the parameters cannot make the output depend on non-fluid regions
    n = 8
    b, coef = synthetic_system(n, 4, fill=0.3)
    p = uo.init_like_library(19).astype(np.float64)
    layout, _, _, _ = uo.param_layout()
    for _, bo, _, cout, _ in layout:
        p[bo:bo + cout] = 0.5  # positive biases: unmasked blocks would be non-zero
    net = uo.Net(n, p)
    net.forward(b, coef)
    m0 = coef != 0
    m1 = np.zeros((4, 4, 4), bool)
    for z in range(4):
        for y in range(4):
            for x in range(4):
                m1[z, y, x] = m0[2 * z:2 * z + 2, 2 * y:2 * y + 2, 2 * x:2 * x + 2].any()
    m2 = np.zeros((2, 2, 2), bool)
    for z in range(2):
        for y in range(2):
            for x in range(2):
                m2[z, y, x] = m1[2 * z:2 * z + 2, 2 * y:2 * y + 2, 2 * x:2 * x + 2].any()
    masks = {0: m0, 1: m1, 2: m2}
    names = ["a0", "b0", "a1", "b1", "a2", "b2", "c1", "d1", "c0", "d0"]
    for i, name in enumerate(names):
        lev = layout[i][4]
        act = net.t[name]
        assert np.all(act[~masks[lev]] == 0.0), name
        assert np.any(act[masks[lev]] > 0.0), name
    assert not m1.all() and m1.any()


def test_bf16_forward_close_to_fp64():
    n = 8
    b, coef = synthetic_system(n, 3)
    p = uo.init_like_library(13)
    z64 = uo.Net(n, p).forward(b, coef)
    z16 = uo.Net(n, p, bf16=True).forward(b, coef)
    assert np.linalg.norm(z16 - z64) <= 3e-2 * np.linalg.norm(z64)


def test_smoothed_net_gradient_matches_finite_differences():
    n = 8
    b, coef = synthetic_system(n, 5)
    p = uo.init_like_library(17).astype(np.float64)
    layout, hw, hb, total = uo.param_layout()
    rng = np.random.default_rng(18)
    for w, bo, cin, cout, _ in layout:
        p[bo:bo + cout] = rng.normal(0.0, 0.05, cout)
    L, g = uo.SmoothedNet(n, p).loss_grad(b, coef)
    idx = [hw + 1, hb] + [int(rng.integers(w, bo)) for w, bo, _, _, _ in layout]
    eps = 1e-6
    for i in idx:
        pp = p.copy()
        pp[i] += eps
        lp = uo.SmoothedNet(n, pp).loss(b, coef)
        pp[i] -= 2 * eps
        lm = uo.SmoothedNet(n, pp).loss(b, coef)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-8, (i, fd, g[i])


def test_smoothed_net_with_zero_network_is_two_jacobi_sweeps():
    # N = 0 (all parameters zero): z = two damped-Jacobi steps from 0
    n = 8
    b, coef = synthetic_system(n, 6)
    net = uo.SmoothedNet(n, np.zeros(uo.param_layout()[3]), omega=0.5)
    z = net.apply(b, coef)
    r = np.where(coef != 0, b.astype(np.float64), 0.0)
    d = np.where(coef != 0, (coef & 7).astype(np.float64), 1.0)
    x1 = 0.5 * r / d
    x2 = x1 + 0.5 * (r - uo.apply_A(coef, x1)) / d
    assert np.allclose(z, np.where(coef != 0, x2, 0.0), rtol=1e-14, atol=1e-14)
