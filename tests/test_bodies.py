"""Rigid bodies from solid particles (Scene::addAllTo -> FreezeParticle ->
AddToRigidBody, src/cpu/UnifiedPhysics.cpp:825-883,1657-1847, finished as
UnifiedPhysics' constructor does at :141-160).

Pins of the oracle: Matrix3x3::CalculateInverse against the reference's own
Matrix3x3.cpp (oracle/_ref); the incremental float accumulation against the
closed forms in float64 (total mass, centre of mass, inertia about it,
velocity, angular momentum about it); bodies against the connected
components of the |x_i - x_j| < h graph (scipy).  The library's host code
must equal the oracle bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as ol
import scenes
import unifiedparticleframeworkcuda_amd as upf
from unifiedparticleframeworkcuda_amd._abi import UPF_ERR_RANGE, ptr


def body_bytes(bodies):
    return b"".join(bytes(b) for b in bodies)


def lattice_block(n3, spacing, origin, rng, jitter=0.0):
    g = np.stack(np.meshgrid(*[np.arange(k) for k in n3], indexing="ij"), -1).reshape(-1, 3).astype(np.float64)
    x = np.asarray(origin) + g * spacing + rng.normal(size=g.shape) * jitter * spacing
    out = np.zeros((len(x), 4), np.float32)
    out[:, :3] = x
    return out


def test_inverse_equals_the_reference_matrix3x3():
    if ol.ref_lib() is None:
        pytest.skip("oracle/_ref not built (no reference tree)")
    rng = np.random.default_rng(1)
    mats = [rng.normal(size=(3, 3)) * 10.0 ** rng.integers(-6, 4) for _ in range(3000)]
    for _ in range(300):  # symmetric positive (inertia-like), some nearly singular
        a = rng.normal(size=(3, 3))
        s = a @ a.T
        s[2] *= 10.0 ** -rng.integers(0, 12)
        mats.append(s)
    mats.append(np.zeros((3, 3)))
    mats.append(np.diag([1.0, 1.0, 0.0]))
    for m in mats:
        m = m.astype(np.float32)
        ok_o, inv_o = ol.inverse3(m)
        ok_r, inv_r = ol.inverse3(m, ref=True)
        assert ok_o == ok_r
        assert np.array_equal(inv_o.view(np.uint32), inv_r.view(np.uint32))


def test_accumulated_body_equals_the_closed_forms():
    rng = np.random.default_rng(2)
    for trial in range(6):
        x = lattice_block((7, 5, 6), 0.01, rng.random(3), rng, jitter=0.2)
        m = (rng.random(len(x)) * 0.002 + 0.0005).astype(np.float32)
        v = np.zeros_like(x)
        v[:, :3] = rng.normal(size=(len(x), 3))
        I, c, vel, L, M = ol.body_accumulate(x, m, v)
        X, V, W = x[:, :3].astype(np.float64), v[:, :3].astype(np.float64), m.astype(np.float64)
        Mx = W.sum()
        com = (W[:, None] * X).sum(0) / Mx
        r = X - com
        Ix = np.einsum("i,ij,ik->jk", W, r, r)
        Ix = np.eye(3) * np.trace(Ix) - Ix
        Lx = (W[:, None] * np.cross(r, V)).sum(0)
        vx = (W[:, None] * V).sum(0) / Mx
        assert abs(M - Mx) <= 1e-5 * Mx
        assert np.allclose(c, com, rtol=0, atol=1e-5 * (np.abs(com).max() + 0.1))
        assert np.allclose(I, Ix, rtol=0, atol=2e-3 * np.abs(Ix).max())
        assert np.allclose(vel, vx, rtol=0, atol=1e-4 * np.abs(V).max())
        assert np.allclose(L, Lx, rtol=0, atol=2e-3 * np.abs(Lx).max())
        assert np.allclose(I, I.T)


def _clusters(rng, p):
    """Solid particles: three lattice blocks (two touching, so they form one
    body), a pair, a lone particle, and a jittered sphere shell."""
    s = float(p.spacing)
    a = lattice_block((6, 4, 5), s, (0.20, 0.20, 0.20), rng, 0.05)
    b = lattice_block((3, 3, 3), s, (0.20 + 6 * s, 0.20, 0.20), rng, 0.05)  # touches a
    c = lattice_block((4, 4, 4), s, (0.50, 0.30, 0.25), rng, 0.05)
    pair = lattice_block((2, 1, 1), s, (0.70, 0.50, 0.40), rng)
    lone = lattice_block((1, 1, 1), s, (0.90, 0.60, 0.60), rng)
    u = rng.normal(size=(300, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    shell = np.zeros((300, 4), np.float32)
    shell[:, :3] = np.asarray([0.4, 0.7, 0.7]) + 0.05 * u
    x = np.concatenate([a, pair, c, lone, b, shell])
    return x[rng.permutation(len(x))]


def test_bodies_are_the_components_of_the_h_graph():
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import connected_components
    p, _, _, _ = scenes.dam_break(8)
    rng = np.random.default_rng(3)
    x = _clusters(rng, p)
    m = np.full(len(x), np.float32(p.initial_mass))
    bid, rel, bodies = ol.bodies_from_particles(p, x, m)
    X = x[:, :3]
    d = X[:, None, :] - X[None, :, :]
    d2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]).astype(np.float32) + d[..., 2] * d[..., 2]
    adj = (d2.astype(np.float32) < np.float32(p.support_radius_sq_limit)) & ~np.eye(len(x), dtype=bool)
    nc, lab = connected_components(csr_matrix(adj), directed=False)
    sizes = np.bincount(lab)
    # bodies: components of two or more, numbered by first particle
    first = {}
    for i, l in enumerate(lab):
        if sizes[l] >= 2 and l not in first:
            first[l] = len(first)
    exp = np.array([first[l] if sizes[l] >= 2 else -1 for l in lab])
    assert np.array_equal(bid, exp)
    assert len(bodies) == len(first) and np.sum(bid < 0) >= 1
    for k, b in enumerate(bodies):
        sel = bid == k
        assert b.mass == pytest.approx(float(m[sel].sum()), rel=1e-5)
        assert np.allclose(rel[sel, :3] + np.asarray(b.pos, np.float32), x[sel, :3], atol=1e-6)
        assert np.all(rel[sel, 3] == m[sel])
        assert list(b.quaternion) == [1, 0, 0, 0] and list(b.linear_momentum) == [0, 0, 0]
    pair_body = [k for k in range(len(bodies)) if np.sum(bid == k) == 2]
    assert pair_body and all(v == 0 for v in bodies[pair_body[0]].inv_inertia_local)  # singular inertia


def test_library_equals_oracle():
    p, _, _, _ = scenes.dam_break(8)
    rng = np.random.default_rng(4)
    for trial in range(4):
        x = _clusters(rng, p)
        m = (rng.random(len(x)) * 0.001 + float(p.initial_mass)).astype(np.float32)
        v = np.zeros_like(x)
        v[:, :3] = rng.normal(size=(len(x), 3)).astype(np.float32)
        vel = None if trial % 2 else v
        got = upf.bodies_from_particles(p, x, m, vel)
        exp = ol.bodies_from_particles(p, x, m, vel)
        assert np.array_equal(got[0], exp[0])
        assert np.array_equal(got[1].view(np.uint32), exp[1].view(np.uint32))
        assert body_bytes(got[2]) == body_bytes(exp[2])


def test_library_equals_oracle_on_obj_models():
    """Solid particles of the procedural OBJ models (the oracle's lattice fill)
    placed apart."""
    import meshes
    import os
    import tempfile
    p, _, _, _ = scenes.dam_break(8)
    parts = []
    with tempfile.TemporaryDirectory() as d:
        for k, text in enumerate((meshes.sphere(0.05, 24, 13), meshes.torus(0.05, 0.02, 24, 12),
                                  meshes.box((0.08, 0.05, 0.06)))):
            path = os.path.join(d, "m%d.obj" % k)
            with open(path, "w") as f:
                f.write(text)
            pts = ol.mesh_fill(ol.obj_read(path), 1.0, float(p.spacing))
            x = np.zeros((len(pts), 4), np.float32)
            x[:, :3] = pts + np.float32([0.2 + 0.25 * k, 0.3, 0.3])
            parts.append(x)
    x = np.concatenate(parts)
    m = np.full(len(x), np.float32(p.initial_mass))
    got = upf.bodies_from_particles(p, x, m)
    exp = ol.bodies_from_particles(p, x, m)
    # one body per model, plus any stray islands the fill's in/out test leaves
    sizes = np.sort(np.bincount(exp[0][exp[0] >= 0]))[::-1]
    assert len(sizes) >= 3 and sizes[:3].sum() >= 0.95 * len(x)
    assert np.array_equal(got[0], exp[0]) and body_bytes(got[2]) == body_bytes(exp[2])
    assert np.array_equal(got[1].view(np.uint32), exp[1].view(np.uint32))


def test_capacity_and_count_only():
    p, _, _, _ = scenes.dam_break(8)
    rng = np.random.default_rng(5)
    x = _clusters(rng, p)
    n = len(x)
    m = np.full(n, np.float32(p.initial_mass))
    L = upf.lib()
    bid = np.zeros(n, np.int32)
    nb = C.c_int()
    assert L.upf_bodies_from_particles(C.byref(p), n, ptr(x), None, ptr(m), ptr(bid, C.c_int), None, None, 0,
                                       C.byref(nb)) == 0
    arr = (upf.Body * 1)()
    assert L.upf_bodies_from_particles(C.byref(p), n, ptr(x), None, ptr(m), ptr(bid, C.c_int), None, arr, 1,
                                       C.byref(nb)) == UPF_ERR_RANGE
    assert nb.value > 1


def test_non_finite_positions_are_rejected():
    p, _, _, _ = scenes.dam_break(8)
    x = np.zeros((3, 4), np.float32)
    x[:, 0] = [0.1, 0.1 + float(p.spacing), np.nan]
    m = np.full(3, np.float32(p.initial_mass))
    with pytest.raises(upf.UPFError):
        upf.bodies_from_particles(p, x, m)
    x[2, 0] = np.inf
    with pytest.raises(upf.UPFError):
        upf.bodies_from_particles(p, x, m)
