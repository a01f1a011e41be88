"""GPU checks of the learned preconditioner's U-Net.

Forward pass (BF16 MFMA convolutions, FP32 accumulation) against the FP64
oracle that rounds every stored activation and weight to bfloat16 as the HIP
path does: 1e-2 normwise relative (bf16 keeps 8 significant bits, and
summation order flips an occasional last bit of an activation).  Parameter
gradients (FP32 backward kernels on the bf16 activations) against the
oracle's FP64 gradient of the same bf16-rounded forward: 4e-2 normwise
relative per layer and cosine >= 0.999 overall.  The same for the smoothed
two-level preconditioner built around the network.  Training: Adam on the
library's own kernels lowers the loss, and the trained network inside
flexible CG reaches the 1e-6 FP64 residual.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import unet_oracle as uo
from tests.unet_cases import synthetic_system
from flip2_amd import scenes


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def biased_params(net, seed):
    p = net.params().astype(np.float64)
    layout, hw, hb, _ = uo.param_layout()
    rng = np.random.default_rng(seed)
    for w, bo, cin, cout, _ in layout:
        p[bo:bo + cout] = rng.normal(0.0, 0.05, cout)
    p[hb] = 0.01
    net.set_params(p.astype(np.float32))
    return net.params()


@pytest.mark.parametrize("n", [8, 16, 24])
def test_forward_matches_oracle(flip2, n):
    net = flip2.UNet(n, seed=5)
    net.set_smoothing(0.0)
    p = biased_params(net, 1 + n)
    b, coef = synthetic_system(n, 4 + n)
    z = net.apply_dense(b, coef).reshape(n, n, n)
    zo = uo.Net(n, p, bf16=True).forward(b, coef)
    assert rel(z, zo) <= 1e-2, rel(z, zo)
    assert np.all(z[coef == 0] == 0.0)
    # and within bf16 precision of the exact FP64 network
    z64 = uo.Net(n, p, bf16=False).forward(b, coef)
    assert rel(z, z64) <= 5e-2


def check_gradients(g, go, tol):
    layout, hw, hb, total = uo.param_layout()
    for i, (w, bo, cin, cout, _) in enumerate(layout):
        assert rel(g[w:bo], go[w:bo]) <= tol, (i, "W", rel(g[w:bo], go[w:bo]))
        assert rel(g[bo:bo + cout], go[bo:bo + cout]) <= tol, (i, "b", rel(g[bo:bo + cout], go[bo:bo + cout]))
    assert rel(g[hw:hw + 16], go[hw:hw + 16]) <= tol
    assert abs(g[hb] - go[hb]) <= tol * max(abs(go[hb]), 1e-6)
    # and the whole gradient points the same way
    cos = float(g.astype(np.float64) @ go / (np.linalg.norm(g) * np.linalg.norm(go)))
    assert cos >= 0.999, cos


# Per-layer tolerance: the GPU backward runs on activations whose bf16 last bit
# can differ from the oracle's (different summation order), and a ReLU whose
# input sits at 0 can switch; layers fed by few voxels (level 1 and 2 at n=16)
# show this most.
GRAD_TOL = 4e-2


@pytest.mark.parametrize("n", [8, 16])
def test_gradient_matches_oracle(flip2, n):
    net = flip2.UNet(n, seed=7)
    net.set_smoothing(0.0)
    p = biased_params(net, 30 + n)
    b, coef = synthetic_system(n, 9 + n)
    loss, g = net.loss_grad(b, coef)
    Lo, go = uo.Net(n, p, bf16=True).loss_grad(b, coef)
    assert abs(loss - Lo) <= 1e-2 * Lo, (loss, Lo)
    check_gradients(g, go, GRAD_TOL)


@pytest.mark.parametrize("n", [8, 16])
def test_smoothed_preconditioner_matches_oracle(flip2, n):
    net = flip2.UNet(n, seed=8)
    net.set_smoothing(2.0 / 3.0)
    p = biased_params(net, 50 + n)
    b, coef = synthetic_system(n, 60 + n)
    z = net.apply_dense(b, coef).reshape(n, n, n)
    ref = uo.SmoothedNet(n, p, omega=2.0 / 3.0, bf16=True)
    zo = ref.apply(b, coef)
    assert rel(z, zo) <= 1e-2, rel(z, zo)
    loss, g = net.loss_grad(b, coef)
    Lo, go = uo.SmoothedNet(n, p, omega=2.0 / 3.0, bf16=True).loss_grad(b, coef)
    assert abs(loss - Lo) <= 1e-2 * Lo, (loss, Lo)
    check_gradients(g, go, GRAD_TOL)


@pytest.mark.parametrize("n", [16, 24])
def test_fused_decoder_matches_stored_activations(flip2, n):
    """The preconditioner apply reads the decoder's concatenated inputs in
    place and runs the output head inside the last convolution; the training
    forward stores every activation.  Both compute the same BF16 values in
    the same order (n=16 also takes the XCD tile remapping, n=24 does not)."""
    net = flip2.UNet(n, seed=11)
    net.set_smoothing(0.0)
    biased_params(net, 70 + n)
    b, coef = synthetic_system(n, 80 + n)
    l_fused, _ = net.loss_grad(b, coef, grad=False)
    l_stored, _ = net.loss_grad(b, coef, grad=True)
    assert abs(l_fused - l_stored) <= 1e-12 * l_stored, (l_fused, l_stored)


def test_sparse_apply_after_the_fluid_moved(flip2):
    """The apply computes only tiles with fluid and clears tiles that the
    previous apply computed: after a system whose fluid fills other tiles,
    the output equals that of a fresh network (same parameters) on the new
    system, bit for bit, and the masked (oracle) network within bf16."""
    n = 32
    net = flip2.UNet(n, seed=12)
    net.set_smoothing(0.0)
    p = biased_params(net, 90)
    b1, coef1 = synthetic_system(n, 91, fill=0.9)   # fluid nearly everywhere
    b2, coef2 = synthetic_system(n, 92, fill=0.25)  # a shallow pool
    net.apply_dense(b1, coef1)
    z2 = net.apply_dense(b2, coef2)
    fresh = flip2.UNet(n, seed=13)
    fresh.set_smoothing(0.0)
    fresh.set_params(p)
    z2f = fresh.apply_dense(b2, coef2)
    assert np.array_equal(z2, z2f)
    zo = uo.Net(n, p, bf16=True).forward(b2, coef2)
    assert rel(z2.reshape(n, n, n), zo) <= 1e-2
    # and back: the tiles of the first system are computed again
    z1 = net.apply_dense(b1, coef1)
    assert np.array_equal(z1, fresh.apply_dense(b1, coef1))


def test_loss_grad_is_repeatable(flip2):
    n = 16
    net = flip2.UNet(n, seed=3)
    b, coef = synthetic_system(n, 21)
    l1, g1 = net.loss_grad(b, coef)
    l2, g2 = net.loss_grad(b, coef)
    # the loss and the weight gradients are sums with atomics: the order may
    # differ in the last bits
    assert abs(l1 - l2) <= 1e-12 * abs(l1)
    assert rel(g1, g2.astype(np.float64)) <= 1e-5


def test_training_and_learned_fcg(flip2):
    """Train on the residuals r_k = b - A x_k that Jacobi-PCG produces in its
    first iterations (k = 0, 2, ..., 64) over 24 dam-break steps; the trained
    network, wrapped in damped-Jacobi smoothing and in brick-local IC(0)
    smoothing, preconditions flexible CG to the 1e-6 FP64 residual in fewer
    iterations than Jacobi-PCG on the next step's system."""
    n = 32
    x, y, z = scenes.dam_break(n, seed=1)
    zero = np.zeros_like(x)
    grid = (0.0, 0.0, 0.0, n, n, n, 1.0 / n)
    sim = flip2.FluidSim(x.size, grid, preconditioner="jacobi")
    sim.set_particles(x, y, z, zero, zero, zero)
    net = flip2.UNet(n, seed=2)
    its = [0, 2, 4, 8, 16, 32, 64]
    net.capture_residuals(sim, its)
    for _ in range(24):
        sim.step()
        assert sim.last_solve()["converged"]
    net.capture_residuals(sim, [])
    nf = net.num_frames()
    assert nf == 24 * len(its)
    hold = 2 * len(its)
    before = net.eval(nf - hold, nf)
    net.train(20, 1e-3, 0, nf - hold, seed=1)
    after = net.eval(nf - hold, nf)
    assert after < 0.25 * before, (before, after)
    for st in ("transfer", "prepare", "system"):
        sim.stage(st)
    sim.set_unet(net)
    sim.set_learned_smoother("jacobi")
    jac = sim.solve_with("jacobi")
    lrn = sim.solve_with("learned")
    assert jac["converged"] and lrn["converged"], (jac, lrn)
    assert lrn["rel_residual"] <= 1e-6
    assert lrn["iterations"] < jac["iterations"], (lrn, jac)
    # the applies replayed from the HIP graph compute what eager launches do
    # (deterministic kernels: same iterations, same residual bit for bit);
    # graph, eager, graph again, with the eager solves between the replays
    sim.set_learned_graph(False)
    eager = sim.solve_with("learned")
    sim.set_learned_graph(True)
    again = sim.solve_with("learned")
    for r in (eager, again):
        assert r["iterations"] == lrn["iterations"] and r["rel_residual"] == lrn["rel_residual"], (r, lrn)
    # residual checks in batches of 8 iterations stop no earlier than checks
    # after every iteration (the default), and reach the same tolerance
    sim.set_flex_check_every(8)
    batched = sim.solve_with("learned")
    sim.set_flex_check_every(1)
    assert batched["converged"] and batched["rel_residual"] <= 1e-6, batched
    assert batched["iterations"] >= lrn["iterations"], (batched, lrn)
    # IC(0) smoothing (the default): converges in about as many iterations
    # or fewer, and the graph replays match eager launches bit for bit too
    sim.set_learned_smoother("ic0")
    ic = sim.solve_with("learned")
    assert ic["converged"] and ic["rel_residual"] <= 1e-6, ic
    assert ic["iterations"] <= lrn["iterations"] + 2, (ic, lrn)
    sim.set_learned_graph(False)
    ic_eager = sim.solve_with("learned")
    sim.set_learned_graph(True)
    assert ic_eager["iterations"] == ic["iterations"] and ic_eager["rel_residual"] == ic["rel_residual"], (ic_eager, ic)
    # the fallback: a cap below what FCG needs hands the solve to Jacobi-PCG
    sim.set_unet(net, fallback_iterations=2)
    fb = sim.solve_with("learned")
    assert fb["used_fallback"] and fb["converged"] and fb["rel_residual"] <= 1e-6, fb
