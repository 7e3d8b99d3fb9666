"""Differentiable PCISPH on the GPU: the rollout equals the FP32 oracle
rollout, the adjoint equals the oracle's adjoint (FP32 and FP64), central
finite differences of the GPU rollout agree with the GPU gradient, the
checkpointed backward pass equals the uncheckpointed one, and the
torch.autograd.Function wrapper returns the same gradient.
"""
import numpy as np
import pytest

import unifiedparticleframeworkcuda_amd as upf
from unifiedparticleframeworkcuda_amd import adjoint as adj
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _loss(x, target, f):
    d = x[f].astype(np.float64) - target[f]
    return float(np.mean(np.sum(d * d, axis=1)))


def _dloss(x, target, f):
    g = np.zeros(x.shape, np.float64)
    g[f] = 2.0 * (x[f].astype(np.float64) - target[f]) / f.sum()
    return g


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_rollout_and_gradient_match_fp32_oracle(gpu):
    p, pos, typ, vel, target, f = adj.differentiable_dam_break(12, 8, 8)
    steps, K = 6, 3
    print("adjoint vs FP32 oracle: %d particles, %d steps, K=%d" % (len(typ), steps, K), flush=True)
    r = adj.Rollout(p, pos, typ, steps, K, checkpoint_every=4)
    x, v = r.forward(vel)
    xo, vo = ol.diff_run(p, pos, vel, typ, steps, K, precision=32)
    # same passes and summation orders; only expf / log1pf of the smooth
    # pressure clamp come from different math libraries
    assert np.abs(x - xo).max() < 2e-6, np.abs(x - xo).max()
    assert np.abs(v - vo).max() < 1e-4 * np.abs(vo).max()
    g = _dloss(x, target, f)
    gx, gv = r.backward(g)
    _, _, ox, ov = ol.diff_run(p, pos, vel, typ, steps, K, gpos=g.astype(np.float32).astype(np.float64),
                               precision=32)
    assert _rel(gv, ov) < 1e-4, _rel(gv, ov)
    assert _rel(gx, ox) < 1e-4, _rel(gx, ox)
    r.close()


def test_checkpointed_backward_equals_full_tape(gpu):
    p, pos, typ, vel, target, f = adj.differentiable_dam_break(10, 8, 8)
    steps, K = 9, 2
    grads = []
    for ck in (1, 4, 9):
        r = adj.Rollout(p, pos, typ, steps, K, checkpoint_every=ck)
        x, _ = r.forward(vel)
        grads.append(r.backward(_dloss(x, target, f)))
        r.close()
    for gx, gv in grads[1:]:
        assert np.array_equal(gv, grads[0][1]) and np.array_equal(gx, grads[0][0])


def test_gradient_matches_fp64_oracle_and_finite_differences(gpu):
    """The gradient workload: 65,536-particle dam break (64x32x32), 64 steps,
    5 iterations, checkpoint every 8 steps, gradient of the mean squared
    distance to a target cloud with respect to the initial velocities."""
    p, pos, typ, vel, target, f = adj.differentiable_dam_break()
    steps, K = 64, 5
    print("adjoint config: %d particles (%d fluid), %d steps" % (len(typ), f.sum(), steps), flush=True)
    r = adj.Rollout(p, pos, typ, steps, K, checkpoint_every=8)
    x, _ = r.forward(vel)
    L0 = _loss(x, target, f)
    _, gv = r.backward(_dloss(x, target, f))
    fwd_ms, bwd_ms, mem = r.timing()
    print("loss %.6g, forward %.2f ms, backward %.2f ms (%.2fx), device %.1f MB" %
          (L0, fwd_ms, bwd_ms, bwd_ms / fwd_ms, mem / 1e6), flush=True)
    # FP64 oracle gradient of the same rollout
    xo, _, _, go = ol.diff_run(p, pos, vel, typ, steps, K, gpos=_dloss(x, target, f))
    e64 = _rel(gv[f], go[f])
    cos = float(np.sum(gv[f] * go[f]) / (np.linalg.norm(gv[f]) * np.linalg.norm(go[f])))
    print("GPU FP32 vs oracle FP64: position rel err %.3g, gradient rel err %.3g, cosine %.6f" %
          (_rel(x[f], xo[f]), e64, cos), flush=True)
    assert e64 < 2e-3 and cos > 0.99999
    # central finite differences on 32 random directions, eps = 1e-3 m/s
    rng = np.random.RandomState(6)
    eps = 1e-3
    fd, an = [], []
    for t in range(32):
        d = np.zeros_like(vel)
        d[f] = rng.normal(size=(int(f.sum()), 3))
        lp = _loss(r.forward(vel + eps * d)[0], target, f)
        lm = _loss(r.forward(vel - eps * d)[0], target, f)
        fd.append((lp - lm) / (2 * eps))
        an.append(float(np.sum(gv.astype(np.float64) * d)))
    fd, an = np.array(fd), np.array(an)
    err = _rel(fd, an)
    print("finite differences: rel err %.3g over 32 directions (max |fd-an|/|an| %.3g)" %
          (err, np.max(np.abs(fd - an) / np.abs(an))), flush=True)
    assert err < 5e-2
    r.close()


def test_torch_function_gives_the_adjoint_gradient(gpu):
    torch = pytest.importorskip("torch")
    p, pos, typ, vel, target, f = adj.differentiable_dam_break(8, 6, 6)
    steps, K = 4, 2
    r = adj.Rollout(p, pos, typ, steps, K, checkpoint_every=2)
    fn = adj.torch_rollout_function()
    v0 = torch.tensor(vel, dtype=torch.float32, requires_grad=True)
    tgt = torch.tensor(target, dtype=torch.float32)
    mask = torch.tensor(f)
    x = fn.apply(v0, r)
    loss = ((x[mask] - tgt[mask]) ** 2).sum(dim=1).mean()
    loss.backward()
    xn, _ = r.forward(vel)
    g = np.zeros_like(xn)
    g[f] = 2.0 * (xn[f] - target[f]) / f.sum()
    _, gv = r.backward(g)
    assert np.allclose(v0.grad.numpy(), gv, rtol=1e-5, atol=1e-12)
    r.close()
