"""The oracle of the differentiable PCISPH rollout and its adjoint.

Pinned by: central finite differences of the FP64 rollout (the textbook check
of a gradient), the closed-form kernels against the reference's own kernel
functions (compiled from its sources into oracle/_ref) and against the
reference tables, and closeness of the closed-form step to the table-based
step of the CPU path restatement.
"""
import ctypes as C

import numpy as np
import pytest

import unifiedparticleframeworkcuda_amd as upf
import oracle_lib as ol


def _scene(n=6, seed=4):
    p, pos, typ, nf = upf.standard_dam_break(n, n, n, seed=0)
    rng = np.random.RandomState(seed)
    vel = np.zeros((len(typ), 3))
    f = typ == upf.UPF_LIQUID
    vel[f] = rng.normal(0.0, 0.1, size=(int(f.sum()), 3))
    return p, pos, typ, vel, f


def _loss(x, target, f):
    d = x[f] - target[f]
    return float(np.mean(np.sum(d * d, axis=1)))


def _dloss(x, target, f):
    g = np.zeros_like(x)
    g[f] = 2.0 * (x[f] - target[f]) / f.sum()
    return g


def test_closed_form_kernels_match_reference_functions():
    h = 0.0200002
    r = np.linspace(0.0, 1.05 * h, 997)
    W, dW, G, dG, L = (np.zeros_like(r) for _ in range(5))
    pd = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    ol.oracle().orc_diff_kernels(h, len(r), pd(r), pd(W), pd(dW), pd(G), pd(dG), pd(L))
    # tables: entry k is the function at k h / 300
    tab = upf.luts(upf.make_params(0.01))
    k = np.arange(300)
    rk = k * h / 300.0
    Wk, dWk, Gk, dGk, Lk = (np.zeros(300) for _ in range(5))
    ol.oracle().orc_diff_kernels(h, 300, pd(rk), pd(Wk), pd(dWk), pd(Gk), pd(dGk), pd(Lk))
    assert np.allclose(Wk, tab[0, :300], rtol=2e-6, atol=1e-3)
    # float tables: (h - r) cancels near h, a few 1e-5 relative there
    assert np.allclose(Gk[1:], tab[1, 1:300], rtol=2e-4)
    assert np.allclose(Lk, tab[2, :300], rtol=2e-4)
    ref = ol.ref_lib()
    if ref is not None:
        rf = r.astype(np.float32)
        m4, pg, vl = (np.zeros(len(r), np.float32) for _ in range(3))
        ref.ref_kernels(h, len(r), ol.ptr(rf), ol.ptr(m4), ol.ptr(pg), ol.ptr(vl))
        inside = r < h * 0.9999
        assert np.allclose(W[inside], m4[inside], rtol=2e-6, atol=1e-2)
        assert np.allclose(G[inside & (r > 0)], pg[inside & (r > 0)], rtol=2e-4)
        assert np.allclose(L[inside], vl[inside], rtol=2e-4)
    # derivatives against central differences of the closed forms
    e = 1e-9
    rr = np.linspace(0.001 * h, 0.999 * h, 401)
    rr = rr[np.abs(rr - 0.5 * h) > 1e-6]
    a, b, c, d_, l1 = (np.zeros_like(rr) for _ in range(5))
    ap, bp, cp_, dp, lp = (np.zeros_like(rr) for _ in range(5))
    am, bm, cm, dm, lm = (np.zeros_like(rr) for _ in range(5))
    ol.oracle().orc_diff_kernels(h, len(rr), pd(rr), pd(a), pd(b), pd(c), pd(d_), pd(l1))
    ol.oracle().orc_diff_kernels(h, len(rr), pd(rr + e), pd(ap), pd(bp), pd(cp_), pd(dp), pd(lp))
    ol.oracle().orc_diff_kernels(h, len(rr), pd(rr - e), pd(am), pd(bm), pd(cm), pd(dm), pd(lm))
    assert np.allclose((ap - am) / (2 * e), b, rtol=1e-5, atol=1e-3 * np.abs(b).max())
    assert np.allclose((cp_ - cm) / (2 * e), d_, rtol=1e-5, atol=1e-6 * np.abs(d_).max())


def test_closed_form_step_tracks_table_step():
    """Same passes as the table-based CPU path: in a compressed block (every
    density well above rest, so the smooth and the hard pressure clamp agree)
    one step with the same fixed iteration count changes the velocities of
    the particles away from the walls alike (the two steps treat the box
    differently).  The tables hold each function at the start of its bin,
    which raises densities by ~15-20 kg/m^3 against the closed form; at
    ~150 kg/m^3 compression that leaves differences of ~10 %, while a dropped
    or sign-flipped term would change the result by O(1)."""
    p = upf.make_params(0.01, 0.001)
    pos, typ, nf = upf.dam_break(p, 14, 14, 14, (0.3, 0.3, 0.3), lattice_rel=0.94)
    upf.density_error_factor(p)
    f = typ == upf.UPF_LIQUID
    margin = 3 * p.support_radius
    inner = f & np.all((pos[:, :3] > np.array(p.collision_min) + margin) &
                       (pos[:, :3] < np.array(p.collision_max) - margin), axis=1)
    assert inner.sum() > 50
    vel = np.zeros((len(typ), 3))
    for K in (1, 3):
        x1, v1 = ol.diff_run(p, pos, vel, typ, 1, K)
        q = p.copy()
        q.min_loops = K
        q.max_loops = K
        o = ol.OracleSim(q, pos, None, typ, precision=64)
        o.step(1, upf.UPF_PCISPH_STATIC)
        r = o.state()  # upload order
        dv_ref = r["vel"][inner, :3].astype(np.float64)
        assert np.abs(v1[inner] - dv_ref).max() < 0.15 * np.abs(dv_ref).max()
        assert np.corrcoef(v1[inner].ravel(), dv_ref.ravel())[0, 1] > 0.98


def test_closed_form_step_conserves_momentum():
    """Without walls, gravity and clamping, every pair force (viscosity,
    pressure) is antisymmetric: total momentum is unchanged by a step."""
    p = upf.make_params(0.01, 0.001)
    pos, typ, nf = upf.dam_break(p, 7, 7, 7, (0.14, 0.14, 0.14), lattice_rel=0.95)
    upf.density_error_factor(p)
    f = typ == upf.UPF_LIQUID
    pos, typ = pos[f], typ[f]
    p.gravity = 0.0
    for a in range(3):
        p.collision_min[a] = 0.0
        p.collision_max[a] = p.vbox_max[a]
    rng = np.random.RandomState(1)
    vel = rng.normal(0, 0.1, size=(len(typ), 3))
    x1, v1 = ol.diff_run(p, pos, vel, typ, 2, 4)
    m0 = vel.sum(axis=0)
    m1 = v1.sum(axis=0)
    assert np.abs(m1 - m0).max() < 1e-9 * len(typ) * (np.abs(v1).max() + 1.0)
    assert np.abs(v1 - vel).max() > 1e-3  # the forces did act


@pytest.mark.parametrize("steps,K", [(1, 1), (3, 3), (6, 4)])
def test_adjoint_matches_finite_differences_fp64(steps, K):
    p, pos, typ, vel, f = _scene(6)
    rng = np.random.RandomState(7)
    x1, v1 = ol.diff_run(p, pos, vel, typ, steps, K)
    target = x1 + rng.normal(0, 0.01, size=x1.shape)
    _, _, gx0, gv0 = ol.diff_run(p, pos, vel, typ, steps, K, gpos=_dloss(x1, target, f))
    assert np.all(gv0[~f] == 0)
    eps = 1e-6
    for t in range(4):
        d = np.zeros_like(vel)
        d[f] = rng.normal(size=(int(f.sum()), 3))
        lp = _loss(ol.diff_run(p, pos, vel + eps * d, typ, steps, K)[0], target, f)
        lm = _loss(ol.diff_run(p, pos, vel - eps * d, typ, steps, K)[0], target, f)
        fd = (lp - lm) / (2 * eps)
        an = float(np.sum(gv0 * d))
        assert abs(fd - an) <= 1e-5 * abs(an) + 1e-12, (t, fd, an)
        # positions: a whole-field step of 1e-6 m already moves some densities
        # across the pressure clamp (a kink); 1e-8 m stays on one side
        ex = 1e-8
        pos2 = pos.astype(np.float64).copy()
        lp = _loss(ol.diff_run(p, pos2[:, :3] + ex * d, vel, typ, steps, K)[0], target, f)
        lm = _loss(ol.diff_run(p, pos2[:, :3] - ex * d, vel, typ, steps, K)[0], target, f)
        fd = (lp - lm) / (2 * ex)
        an = float(np.sum(gx0 * d))
        assert abs(fd - an) <= 1e-4 * abs(an) + 1e-10, ("x", t, fd, an)


def test_adjoint_of_final_velocity_term():
    p, pos, typ, vel, f = _scene(5)
    rng = np.random.RandomState(3)
    steps, K = 3, 2
    gvel = np.zeros_like(vel)
    gvel[f] = rng.normal(size=(int(f.sum()), 3))
    gpos = np.zeros_like(vel)
    _, _, _, gv0 = ol.diff_run(p, pos, vel, typ, steps, K, gpos=gpos, gvel=gvel)
    d = np.zeros_like(vel)
    d[f] = rng.normal(size=(int(f.sum()), 3))
    eps = 1e-6
    vp = ol.diff_run(p, pos, vel + eps * d, typ, steps, K)[1]
    vm = ol.diff_run(p, pos, vel - eps * d, typ, steps, K)[1]
    fd = float(np.sum(gvel * (vp - vm))) / (2 * eps)
    an = float(np.sum(gv0 * d))
    assert abs(fd - an) <= 1e-5 * abs(an)
