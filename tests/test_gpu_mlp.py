"""GPU tests of the learned warm start: MFMA inference and gradients against the
numpy reference (tests/mlp_reference.py), training, and the solver fed by the
model against the oracle fed the same initial impulses.

Tolerances.  The device accumulates bf16 products in fp32 in MFMA order, the
reference in float64; a different summation order moves a hidden activation
across a bf16 rounding boundary now and then.  Measured on the CPU between two
summation orders: mean error 2e-6, max 4.5e-2 of the output scale.  The tests
allow mean 1e-4, 99.9th percentile 1e-2 and max 0.1 of the output scale; a
wrong operand layout gives errors of order 1 everywhere.
"""
import numpy as np
import pytest

import jbgpu
import mlp_reference as mr
import pyoracle
from jbgpu import scenes

pytestmark = pytest.mark.gpu


def _model(seed, last_scale=1.0):
    rng = np.random.default_rng(seed)
    return mr.init_params(rng, last_scale), mr.random_norm(rng), rng


def test_mlp_params_round_trip():
    p, norm, _ = _model(1)
    ctx = jbgpu.Context()
    with pytest.raises(jbgpu.JbgpuError):
        ctx.mlp_infer(np.zeros((4, 16), np.float32))  # no model yet
    ctx.mlp_set_params(p, norm)
    w, nm = ctx.mlp_get_params()
    assert np.array_equal(w, p) and np.array_equal(nm, norm)
    bad = norm.copy()
    bad[33] = 0.0
    with pytest.raises(jbgpu.JbgpuError):
        ctx.mlp_set_params(p, bad)
    ctx.close()


@pytest.mark.parametrize("n", [1, 15, 16, 17, 5003, 70000])
def test_mlp_inference_matches_reference(n):
    p, norm, rng = _model(5)
    x = (rng.normal(size=(n, 16)) * 2).astype(np.float32)
    ctx = jbgpu.Context()
    ctx.mlp_set_params(p, norm)
    y = ctx.mlp_infer(x)
    ctx.close()
    yr = mr.forward(p, norm, x)
    e = np.abs(y - yr) / norm[32:35]
    print(f"[mlp infer n={n}] mean {e.mean():.2e} p99.9 {np.quantile(e, 0.999):.2e} max {e.max():.2e}", flush=True)
    assert e.mean() < 1e-4 and np.quantile(e, 0.999) < 1e-2 and e.max() < 0.1


def test_mlp_gradient_matches_reference():
    p, norm, rng = _model(7)
    n = 3001
    x = (rng.normal(size=(n, 16)) * 1.5).astype(np.float32)
    t = rng.normal(size=(n, 3)).astype(np.float32) * norm[32:35] + norm[35:38]
    ctx = jbgpu.Context()
    ctx.mlp_set_params(p, norm)
    g, loss = ctx.mlp_grad(x, t)
    ctx.close()
    gr, loss_r = mr.grad(p, norm, x, t)
    assert abs(loss - loss_r) <= 1e-3 * loss_r
    o = 0
    for name, shp in mr.SIZES:
        k = int(np.prod(shp))
        a, b = g[o:o + k].astype(np.float64), gr[o:o + k]
        cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b))
        rel = np.abs(a - b).max() / np.abs(b).max()
        print(f"[mlp grad {name}] cos {cos:.7f} max rel {rel:.2e}", flush=True)
        assert cos > 0.9999 and rel < 2e-2, name
        o += k
    # against the exact float64 gradient of the unrounded model
    nn = norm.astype(np.float64)
    ge = mr.exact_grad(p.astype(np.float64), (x - nn[:16]) * nn[16:32], (t - nn[35:38]) / nn[32:35])
    assert g.astype(np.float64) @ ge / (np.linalg.norm(g) * np.linalg.norm(ge)) > 0.999


def test_mlp_training_fits_a_teacher():
    pt, norm, rng = _model(9)
    n = 40000
    x = (rng.normal(size=(n, 16)) / norm[16:32] + norm[:16]).astype(np.float32)
    t = mr.forward(pt, norm, x).astype(np.float32)
    ctx = jbgpu.Context()
    losses = ctx.mlp_train(x, t, epochs=30, batch=2048, lr=2e-3, seed=3)
    print(f"[mlp train] losses {np.array2string(losses, precision=4)}", flush=True)
    assert np.all(np.isfinite(losses)) and losses[-1] < 0.25 * losses[0]
    assert np.all(np.diff(losses[::5]) < 0)  # still improving every five epochs
    w, nm = ctx.mlp_get_params()
    y = ctx.mlp_infer(x[:4096])
    ctx.close()
    # the trained model on the device agrees with the reference evaluation of the same weights
    yr = mr.forward(w, nm, x[:4096])
    assert np.abs(y - yr).max() / nm[32:35].min() < 0.1
    assert np.mean((y - t[:4096]) ** 2) < np.var(t[:4096], axis=0).mean() * 0.25


def test_learned_warm_start_solver_parity():
    """The world stepped with model warm starts equals the oracle fed the model's impulses, bit for bit."""
    sc = scenes.boxes_grid(seed=4, nx=5, ny=4, nz=5, base=0.5, spacing=1.0)
    shapes, verts, bodies = scenes.build_records(sc, jbgpu)
    ctx = jbgpu.Context(num_iterations=10)
    ctx.world_create(shapes, verts, bodies)
    orc = pyoracle.World(shapes, verts, bodies, iterations=10, order=jbgpu.ORDER_EXPLICIT)
    p, _, _ = _model(11, last_scale=0.05)
    norm = np.concatenate([np.zeros(16), np.ones(16), [0.5, 0.05, 0.05], [0.1, 0.0, 0.0]]).astype(np.float32)
    ctx.mlp_set_params(p, norm)
    ctx.solver_use_mlp(True)
    ctx.solver_enable_features(True)
    orc.enable_features(True)
    checked = 0
    for step in range(40):
        b0 = ctx.bodies()
        prox, s0 = ctx.bp_proxies()
        k0, m0 = ctx.pairs()
        ctx.step()
        order, _ = ctx.solver_order()
        na = len(order)
        if na == 0:
            continue
        feats = ctx.solver_features(4 * na)
        _, man = ctx.pairs()
        mm = man[order]
        npts = mm[:, 0].view(np.int32)
        dist = mm[:, 4:].reshape(na, 4, 32)[:, :, 15]
        valid = (np.arange(4)[None, :] < npts[:, None]) & (dist <= 0)
        valid = valid.reshape(-1)
        pred = ctx.mlp_infer(np.where(valid[:, None], feats, 0.0))
        orc.set_bodies(b0)
        orc.set_proxies(prox, s0)
        orc.set_pairs(k0, m0)
        orc.set_order(order)
        orc.set_learned_init(pred[valid])
        orc.step()
        fo = orc.features()
        assert np.array_equal(fo.view(np.uint32), feats[valid].view(np.uint32)), f"features step {step}"
        bg, bo = ctx.bodies(), orc.bodies()
        assert np.array_equal(bg.view(np.uint32), bo.view(np.uint32)), f"bodies step {step}"
        checked += int(valid.sum())
    print(f"[learned parity] {checked} contacts checked", flush=True)
    assert checked > 1000
    ctx.close()


def test_learned_warm_start_parity_on_the_pyramid_scene():
    """The learned warm-start configuration at its full size: 64k boxes in 32 pyramids and walls plus 2,000
    falling spheres.  60 device steps with the model driving the warm start, then two steps checked bit for bit
    against the oracle fed the model's impulses (features bit-exact too)."""
    sc = scenes.pyramids(seed=1)
    shapes, verts, bodies = scenes.build_records(sc, jbgpu)
    print(f"[learned C5] {sc.n_bodies} bodies", flush=True)
    ctx = jbgpu.Context(num_iterations=10)
    ctx.world_create(shapes, verts, bodies)
    orc = pyoracle.World(shapes, verts, bodies, iterations=10, order=jbgpu.ORDER_EXPLICIT)
    p, _, _ = _model(13, last_scale=0.05)
    norm = np.concatenate([np.zeros(16), np.ones(16), [0.5, 0.05, 0.05], [0.1, 0.0, 0.0]]).astype(np.float32)
    ctx.mlp_set_params(p, norm)
    ctx.solver_use_mlp(True)
    ctx.step(n=60)
    ctx.solver_enable_features(True)
    orc.enable_features(True)
    for step in range(2):
        b0 = ctx.bodies()
        prox, s0 = ctx.bp_proxies()
        k0, m0 = ctx.pairs()
        ctx.step()
        order, _ = ctx.solver_order()
        na = len(order)
        feats = ctx.solver_features(4 * na)
        _, man = ctx.pairs()
        mm = man[order]
        npts = mm[:, 0].view(np.int32)
        dist = mm[:, 4:].reshape(na, 4, 32)[:, :, 15]
        valid = ((np.arange(4)[None, :] < npts[:, None]) & (dist <= 0)).reshape(-1)
        pred = ctx.mlp_infer(np.where(valid[:, None], feats, 0.0))
        orc.set_bodies(b0)
        orc.set_proxies(prox, s0)
        orc.set_pairs(k0, m0)
        orc.set_order(order)
        orc.set_learned_init(pred[valid])
        orc.step()
        fo = orc.features()
        assert np.array_equal(fo.view(np.uint32), feats[valid].view(np.uint32)), f"features step {61 + step}"
        bg, bo = ctx.bodies(), orc.bodies()
        assert np.array_equal(bg.view(np.uint32), bo.view(np.uint32)), f"bodies step {61 + step}"
        print(f"[learned C5] step {61 + step}: {int(valid.sum())} contacts, bit-exact", flush=True)
    ctx.close()


def test_trained_model_reduces_solver_iterations():
    """What the trained network gains over the plain warm start: trained on the device (Adam, own forward and
    backward kernels) on converged solves of seeds 1000-1002 of the pyramid scene (reduced to 8,000 boxes in 8
    pyramids and walls plus 250 spheres to keep the test short), then the solver warm-started by it runs fewer
    iterations to a residual of 1e-2 than with 0.85 x the cached impulses on seed 1, and its residual after 20
    iterations is lower.  The solves are capped at 100 iterations; neither warm start reaches 1e-3 there."""
    from jbgpu import learned
    kw = dict(n_boxes=8000, n_pyramids=8, n_spheres=250, area=80.0)
    x, y = learned.collect(range(1000, 1003), kw, settle=60, samples=4, every=5)
    ctx = jbgpu.Context()
    losses = ctx.mlp_train(x, y, epochs=40, batch=4096, lr=2e-3, seed=0)
    params, norm = ctx.mlp_get_params()
    ctx.close()
    ev = learned.evaluate(params, norm, seed=1, scene_kw=kw, settle=60, steps=20, modes=("plain", "learned"))
    pl, le = ev["plain"], ev["learned"]
    print(f"[learned gain] {len(x)} training contacts, loss {losses[0]:.3f} -> {losses[-1]:.3f}; iterations to 1e-2: "
          f"plain {pl['iterations_to']['0.01']:.2f}, learned {le['iterations_to']['0.01']:.2f}; residual after 20: "
          f"plain {pl['residual_after_20']:.4e}, learned {le['residual_after_20']:.4e}", flush=True)
    assert losses[-1] < losses[0]
    assert le["iterations_to"]["0.01"] < pl["iterations_to"]["0.01"]
    assert le["residual_after_20"] < pl["residual_after_20"]
