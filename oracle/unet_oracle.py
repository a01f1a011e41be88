"""FP64 restatement of the learned preconditioner's U-Net (test infrastructure).

The network has no counterpart in the reference (its pressure solve is the
brick-local SOR of particles.cu:calcPressuresGS); this file restates the HIP
implementation's own definition so its forward pass, parameter gradients and
training loss can be checked:

  input  x[...,0] = bf16(b / s), x[...,1] = fluid mask, channels 2..15 zero,
         s = RMS of b over fluid cells (FP32)
  masks  M0 = fluid cells, M1 = any(M0) over 2x2x2 children, M2 likewise
  a0 = M0 relu(conv(x, 16->16))      b0 = M0 relu(conv(a0, 16->16))        level 0
  a1 = M1 relu(conv(pool(b0), 16->32))  b1 = M1 relu(conv(a1, 32->32))     level 1
  a2 = M2 relu(conv(pool(b1), 32->64))  b2 = M2 relu(conv(a2, 64->64))     level 2
  c1 = M1 relu(conv([up(b2), b1], 96->32))  d1 = M1 relu(conv(c1, 32->32))
  c0 = M0 relu(conv([up(d1), b0], 48->16))  d0 = M0 relu(conv(c0, 16->16))
  z  = s * mask * (d0 . w + beta)
  loss = ||A z - b||^2 / ||b||^2 over fluid cells

Every activation is zero outside its level's mask, so the network's output
does not depend on anything outside the fluid and blocks of the grid with no
fluid need not be computed at all (the HIP apply skips them).

conv is a 3x3x3 cross-correlation with zero padding, weights W[tap][cout][cin]
with tap = (dz*3 + dy)*3 + dx for offsets (dz-1, dy-1, dx-1); pool is the mean
of 2x2x2 children, up is nearest-neighbour upsampling. Arrays are [z][y][x][C].
A is the 7-point operator of the solver: (A x)_v = diag_v x_v - sum of x over
the fluid neighbours flagged in coef_v (bits 3..8: -x, +x, -y, +y, -z, +z).

With `bf16=True` every stored activation is rounded to bfloat16 (nearest even)
and the weights used by the convolutions are the bf16 copies, as in the HIP
forward pass; the backward pass is the exact FP64 derivative of the forward
function evaluated at those (rounded) activations.

Parity status: conv3 is pinned against scipy.ndimage.correlate, pool/up
against numpy reshapes, the backward pass against central finite differences
of the FP64 forward loss, and the operator A against the textbook 7-point
Laplacian (tests/test_oracle_unet.py).
"""
from __future__ import annotations

import numpy as np

LAYERS = [(16, 16, 0), (16, 16, 0), (16, 32, 1), (32, 32, 1), (32, 64, 2),
          (64, 64, 2), (96, 32, 1), (32, 32, 1), (48, 16, 0), (16, 16, 0)]


def param_layout():
    """[(w_offset, b_offset, cin, cout, level)], head_w offset, head_b offset, total."""
    out, off = [], 0
    for cin, cout, lev in LAYERS:
        w = off
        off += 27 * cin * cout
        out.append((w, off, cin, cout, lev))
        off += cout
    hw = off
    off += 16
    hb = off
    off += 1
    return out, hw, hb, off


def bf16_round(a):
    """Round to bfloat16 (nearest, ties to even) and return float64."""
    f = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)


def conv3(x, W, bias):
    """x [n,n,n,cin], W [27,cout,cin], bias [cout] -> [n,n,n,cout]."""
    n = x.shape[0]
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    y = np.broadcast_to(np.asarray(bias, np.float64), x.shape[:3] + (W.shape[1],)).copy()
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                tap = (dz * 3 + dy) * 3 + dx
                y += xp[dz:dz + n, dy:dy + n, dx:dx + n, :] @ W[tap].T
    return y


def conv3_backward(x, W, g):
    """Gradients of sum(g * conv3(x, W, b)): (dx, dW, db)."""
    n = x.shape[0]
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    dxp = np.zeros_like(xp)
    dW = np.zeros_like(W)
    gf = g.reshape(-1, g.shape[-1])
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                tap = (dz * 3 + dy) * 3 + dx
                xs = xp[dz:dz + n, dy:dy + n, dx:dx + n, :].reshape(-1, x.shape[-1])
                dW[tap] = gf.T @ xs
                dxp[dz:dz + n, dy:dy + n, dx:dx + n, :] += g @ W[tap]
    return dxp[1:-1, 1:-1, 1:-1, :], dW, gf.sum(axis=0)


def pool(x):
    n = x.shape[0] // 2
    return x.reshape(n, 2, n, 2, n, 2, x.shape[-1]).mean(axis=(1, 3, 5))


def pool_backward(g):
    return np.repeat(np.repeat(np.repeat(g, 2, 0), 2, 1), 2, 2) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)


def pool_any(m):
    """Coarse mask: true where any of the 2x2x2 children is."""
    n = m.shape[0] // 2
    return np.asarray(m, bool).reshape(n, 2, n, 2, n, 2).any(axis=(1, 3, 5))


def level_masks(mask):
    m1 = pool_any(mask)
    return [np.asarray(mask, bool), m1, pool_any(m1)]


def up_backward(g):
    n = g.shape[0] // 2
    return g.reshape(n, 2, n, 2, n, 2, g.shape[-1]).sum(axis=(1, 3, 5))


def apply_A(coef, x):
    """coef, x: [n,n,n] (u16 operator coefficients, values)."""
    coef = np.asarray(coef).astype(np.int64)
    x = np.asarray(x, np.float64)
    diag = (coef & 7).astype(np.float64)
    m = coef >> 3
    xp = np.pad(x, 1)
    n = x.shape[0]
    c = slice(1, n + 1)
    nb = [xp[c, c, 0:n], xp[c, c, 2:n + 2], xp[c, 0:n, c], xp[c, 2:n + 2, c], xp[0:n, c, c], xp[2:n + 2, c, c]]
    y = diag * x
    for d in range(6):
        y -= np.where((m >> d) & 1, nb[d], 0.0)
    return np.where(coef != 0, y, 0.0)


def make_input(b, coef):
    """(x0 [n,n,n,16], mask [n,n,n], s) as the HIP input assembly."""
    b32 = np.asarray(b, np.float32)
    mask = np.asarray(coef) != 0
    cnt = int(mask.sum())
    ss = float((b32.astype(np.float64)[mask] ** 2).sum())
    s = np.sqrt(ss / (cnt if cnt else 1))
    s32 = np.float32(s if s > 0 else 1.0)
    x0 = np.zeros(b32.shape + (16,), np.float64)
    x0[..., 0] = np.where(mask, bf16_round(b32 / s32), 0.0)
    x0[..., 1] = mask.astype(np.float64)
    return x0, mask, float(s32)


class Net:
    """Forward/backward of the U-Net in FP64 for a flat parameter vector."""

    def __init__(self, n: int, params, bf16: bool = False):
        self.n = n
        self.layout, self.hw, self.hb, self.total = param_layout()
        p = np.asarray(params, np.float64)
        assert p.size == self.total
        self.p = p
        self.bf16 = bf16

    def _W(self, i):
        w, b, cin, cout, _ = self.layout[i]
        W = self.p[w:w + 27 * cin * cout].reshape(27, cout, cin)
        if self.bf16:
            W = bf16_round(W)
        return W, self.p[b:b + cout]

    def _act(self, a):
        return bf16_round(a) if self.bf16 else a

    def forward(self, b, coef):
        x0, mask, s = make_input(b, coef)
        masks = level_masks(mask)
        t = {"x0": x0}
        order = [("x0", "a0", 0), ("a0", "b0", 1), ("p0", "a1", 2), ("a1", "b1", 3), ("p1", "a2", 4),
                 ("a2", "b2", 5), ("u1", "c1", 6), ("c1", "d1", 7), ("u0", "c0", 8), ("c0", "d0", 9)]
        for src, dst, i in order:
            if src == "p0":
                t["p0"] = self._act(pool(t["b0"]))
            if src == "p1":
                t["p1"] = self._act(pool(t["b1"]))
            if src == "u1":
                t["u1"] = np.concatenate([up(t["b2"]), t["b1"]], axis=-1)
            if src == "u0":
                t["u0"] = np.concatenate([up(t["d1"]), t["b0"]], axis=-1)
            W, bias = self._W(i)
            lev = self.layout[i][4]
            t[dst] = self._act(np.maximum(conv3(t[src], W, bias), 0.0) * masks[lev][..., None])
        hw = self.p[self.hw:self.hw + 16]
        hb = self.p[self.hb]
        z = np.where(mask, s * (t["d0"] @ hw + hb), 0.0)
        self.t, self.mask, self.s = t, mask, s
        return z

    def loss(self, b, coef):
        z = self.forward(b, coef)
        b64 = np.where(self.mask, np.asarray(b, np.float32).astype(np.float64), 0.0)
        res = np.where(self.mask, apply_A(coef, z) - b64, 0.0)
        bb = float((b64 ** 2).sum())
        self.res, self.bb = res, bb
        return float((res ** 2).sum()) / bb if bb > 0 else 0.0

    def loss_grad(self, b, coef):
        """(loss, gradient w.r.t. the flat parameter vector)."""
        L = self.loss(b, coef)
        if self.bb <= 0:
            return L, np.zeros(self.total)
        return L, self._backward(2.0 * apply_A(coef, self.res) / self.bb)

    def _backward(self, dz):
        """Parameter gradient for dL/dz (z = the last forward's output)."""
        g = np.zeros(self.total)
        t = self.t
        dh = np.where(self.mask, dz * self.s, 0.0)
        g[self.hw:self.hw + 16] = (t["d0"] * dh[..., None]).reshape(-1, 16).sum(axis=0)
        g[self.hb] = dh.sum()
        hw = self.p[self.hw:self.hw + 16]
        grad = {"d0": dh[..., None] * hw}

        def back(i, src, dst):
            w, bo, cin, cout, _ = self.layout[i]
            gy = grad[dst] * (t[dst] > 0)
            W, _ = self._W(i)
            dx, dW, db = conv3_backward(t[src], W, gy)
            g[w:w + 27 * cin * cout] = dW.reshape(-1)
            g[bo:bo + cout] = db
            grad[src] = grad.get(src, 0.0) + dx

        back(9, "c0", "d0")
        back(8, "u0", "c0")
        grad["d1"] = up_backward(grad["u0"][..., :32])
        grad["b0"] = grad["u0"][..., 32:].copy()
        back(7, "c1", "d1")
        back(6, "u1", "c1")
        grad["b2"] = up_backward(grad["u1"][..., :64])
        grad["b1"] = grad["u1"][..., 64:].copy()
        back(5, "a2", "b2")
        back(4, "p1", "a2")
        grad["b1"] = grad["b1"] + pool_backward(grad["p1"])
        back(3, "a1", "b1")
        back(2, "p0", "a1")
        grad["b0"] = grad["b0"] + pool_backward(grad["p0"])
        back(1, "a0", "b0")
        back(0, "x0", "a0")
        return g


def init_like_library(seed: int = 1234):
    """Not the library's generator: a seeded normal init with the same scales
    (He normal for convs, 0.1 for the head, zero biases)."""
    layout, hw, hb, total = param_layout()
    rng = np.random.default_rng(seed)
    p = np.zeros(total)
    for w, _, cin, cout, _ in layout:
        p[w:w + 27 * cin * cout] = rng.normal(0.0, np.sqrt(2.0 / (27 * cin)), 27 * cin * cout)
    p[hw:hw + 16] = rng.normal(0.0, 0.1, 16)
    return p.astype(np.float32)


def jacobi(coef, r, omega):
    """omega D^-1 r on fluid cells (D = operator diagonal)."""
    coef = np.asarray(coef).astype(np.int64)
    d = (coef & 7).astype(np.float64)
    return np.where(coef != 0, omega * np.asarray(r, np.float64) / np.where(d > 0, d, 1.0), 0.0)


class SmoothedNet(Net):
    """Two-level preconditioner around the network N:
         z1 = w D^-1 r,  r1 = r - A z1,  y = N(r1),
         z  = (z1 + y) + w D^-1 (r - A (z1 + y))
    i.e. a damped-Jacobi pre- and post-smoothing step with the network as the
    coarse correction.  loss = ||A z - r||^2 / ||r||^2; the gradient reaches
    the network through dL/dy = (I - w A D^-1) dL/dz (A symmetric)."""

    def __init__(self, n, params, omega=2.0 / 3.0, bf16=False):
        super().__init__(n, params, bf16)
        self.omega = omega

    def apply(self, r, coef):
        coef = np.asarray(coef)
        r64 = np.where(coef != 0, np.asarray(r, np.float32).astype(np.float64), 0.0)
        z1 = jacobi(coef, r64, self.omega)
        r1 = np.where(coef != 0, r64 - apply_A(coef, z1), 0.0)
        y = self.forward(r1.astype(np.float32), coef)
        x = z1 + y
        z = x + jacobi(coef, r64 - apply_A(coef, x), self.omega)
        self.r64 = r64
        return np.where(coef != 0, z, 0.0)

    def loss(self, r, coef):
        z = self.apply(r, coef)
        res = np.where(np.asarray(coef) != 0, apply_A(coef, z) - self.r64, 0.0)
        rr = float((self.r64 ** 2).sum())
        self.res, self.bb = res, rr
        return float((res ** 2).sum()) / rr if rr > 0 else 0.0

    def loss_grad(self, r, coef):
        L = self.loss(r, coef)
        g = np.zeros(self.total)
        if self.bb <= 0:
            return L, g
        gz = 2.0 * apply_A(coef, self.res) / self.bb
        gy = gz - apply_A(coef, jacobi(coef, gz, self.omega))
        return L, self._backward(gy)
