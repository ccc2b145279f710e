"""The no-allocation entry point jbgpu_mlp_infer_static against jbgpu_mlp_infer (the entry point it mirrors):
identical predictions on two calls in a row, the second smaller, and the size limit."""
import numpy as np
import pytest

import jbgpu
import mlp_reference as mr

pytestmark = pytest.mark.gpu


def test_static_inference_matches_allocating_inference():
    rng = np.random.default_rng(21)
    p, norm = mr.init_params(rng, 1.0), mr.random_norm(rng)
    ctx = jbgpu.Context()
    ctx.mlp_set_params(p, norm)
    for n in (jbgpu.MLP_STATIC_MAX, 1237):  # the full buffer, then a smaller call over the last one's contents
        x = (rng.normal(size=(n, 16)) * 2).astype(np.float32)
        want = ctx.mlp_infer(x)
        got = ctx.mlp_infer_static(x)
        print(f"[static infer] n={n}", flush=True)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), n
    with pytest.raises(jbgpu.JbgpuError, match=r"\(-4\).*at most 65536"):
        ctx.mlp_infer_static(np.zeros((jbgpu.MLP_STATIC_MAX + 1, 16), np.float32))
    assert len(ctx.mlp_infer_static(np.zeros((0, 16), np.float32))) == 0
    ctx.close()


def test_static_inference_needs_a_model():
    ctx = jbgpu.Context()
    with pytest.raises(jbgpu.JbgpuError, match=r"\(-3\)"):
        ctx.mlp_infer_static(np.zeros((4, 16), np.float32))
    ctx.close()
