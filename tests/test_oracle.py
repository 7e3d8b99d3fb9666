"""Pins the CPU oracle before it is trusted as the checker.

* against the reference's own translation units compiled into oracle/_ref
  (constants, kernel tables, z-table, BoxQuery, rigid-body update) when the
  reference tree is present in this container;
* against closed forms and brute force (kernel normalisation, Morton bit
  interleave, neighbour sets, stable sort, the density error factor of an
  ideal lattice, conservation laws of the step).
The reference ships no tests, fixtures or golden vectors for this path, so
these are the pins; tests/golden/ holds vectors written from oracle/_ref for
machines where the reference tree is absent.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle_lib as ol
import unifiedparticleframeworkcuda_amd as upf
from unifiedparticleframeworkcuda_amd._abi import ptr, Params

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
# the reference units are loaded by the tests that use them, never at import
needs_ref = pytest.mark.skipif(not ol.ref_built(), reason="oracle/_ref not built (no reference tree)")


def orc_luts(h):
    out = np.zeros(1505, np.float32)
    ks = np.zeros(1, np.float32)
    ol.oracle().orc_luts(h, ptr(out), ptr(ks))
    return out, ks[0]


# -------------------------------------------------------------- constants
@needs_ref
@pytest.mark.parametrize("spacing", [0.01, 0.025, 0.02, 0.005])
def test_constants_match_reference(spacing):
    ref = np.zeros(32, np.float32)
    m = ol.ref_lib().ref_constants(spacing, ptr(ref))
    p = Params()
    ol.oracle().orc_params_init(spacing, C.byref(p))
    mine = [p.spacing, p.particle_radius, p.support_radius, p.force_distance, p.initial_mass, p.rest_density,
            p.gravity, p.visc_const, 0.0012, p.gas_const_wcsph, p.density_error_factor_parameter,
            p.max_density_error_allowed, p.min_loops, p.max_loops, p.spring_coefficient,
            p.spring_coefficient_boundary, p.damping_coefficient, p.terminal_speed,
            p.surface_tension_gamma, p.surface_adhesion_beta]
    assert np.array_equal(np.float32(mine), ref[:20])


def test_product_params_equal_oracle_params():
    for spacing in (0.01, 0.025, 0.0075):
        a = upf.make_params(spacing)
        b = Params()
        ol.oracle().orc_params_init(spacing, C.byref(b))
        for name in ("spacing", "particle_radius", "support_radius", "force_distance", "initial_mass",
                     "rest_density", "visc_const", "gravity", "dt", "gas_const_wcsph", "density_error_factor_parameter",
                     "max_density_error_allowed", "min_loops", "max_loops", "max_boundary_force"):
            assert getattr(a, name) == getattr(b, name), name


# ------------------------------------------------------------------ LUTs
@needs_ref
@pytest.mark.parametrize("h", [0.0200002, 0.05000050, 0.0100001, 0.0400004])
def test_luts_bit_equal_reference(h):
    h = float(np.float32(h))
    ref = np.zeros(1505, np.float32)
    ks = np.zeros(1, np.float32)
    d = np.zeros(1, np.float32)
    z = np.zeros(1, np.float32)
    ol.ref_lib().ref_luts(h, ptr(ref), ptr(ks), 0, ptr(d), ptr(z), ptr(z), ptr(z))
    mine, ks2 = orc_luts(h)
    assert np.array_equal(mine.view(np.uint32), ref.view(np.uint32))
    assert ks2 == ks[0]
    prod = upf.luts(upf.make_params(h / 2.00002)).ravel()
    p = upf.make_params(h / 2.00002)
    if p.support_radius == np.float32(h):
        assert np.array_equal(prod.view(np.uint32), ref.view(np.uint32))


@needs_ref
def test_analytic_kernels_and_lookups_match_reference():
    h = np.float32(0.0200002)
    rng = np.random.RandomState(0)
    d = np.concatenate([rng.uniform(0, 1.05 * h, 20000), np.linspace(0, h, 301) * (1 + 1e-7)]).astype(np.float32)
    w, g, v = (np.zeros_like(d) for _ in range(3))
    ref = np.zeros(1505, np.float32)
    ks = np.zeros(1, np.float32)
    ol.ref_lib().ref_luts(h, ptr(ref), ptr(ks), len(d), ptr(d), ptr(w), ptr(g), ptr(v))
    idx = np.zeros(len(d), np.int32)
    w2, g2, v2 = (np.zeros_like(d) for _ in range(3))
    ol.oracle().orc_lut_lookup(h, len(d), ptr(d), ptr(idx, C.c_int), ptr(w2), ptr(g2), ptr(v2))
    assert np.array_equal(w, w2) and np.array_equal(g, g2) and np.array_equal(v, v2)


def test_m4_closed_forms():
    """KernelM4: W(0) = 8/(pi h^3), W(h) = 0, integral over the ball = 1."""
    h = np.float32(0.0200002)
    lut, ks = orc_luts(h)
    assert abs(ks - 8.0 / (np.pi * float(h) ** 3)) / ks < 1e-6
    assert lut[300] == 0.0 and lut[301 + 300] == 0.0
    r = (np.arange(300) + 0.5) / 300 * float(h)
    # analytic M4 at bin midpoints integrated over the ball
    s = r / float(h)
    W = 8 / (np.pi * float(h) ** 3) * np.where(s < 0.5, 1 - 6 * s * s + 6 * s ** 3, 2 * (1 - s) ** 3)
    assert abs(np.sum(W * 4 * np.pi * r * r) * float(h) / 300 - 1.0) < 1e-4
    # the table samples the bin's lower edge
    s0 = np.arange(300) / 300.0
    W0 = 8 / (np.pi * float(h) ** 3) * np.where(s0 < 0.5, 1 - 6 * s0 * s0 + 6 * s0 ** 3, 2 * (1 - s0) ** 3)
    assert np.allclose(lut[:300], W0, rtol=1e-5, atol=1e-6 * W0.max())  # float (1 - s)^3 near h
    # pressure gradient table: -45/(pi h^6) (h-r)^2 / r, zero at r = 0
    pg = lut[301:601]
    r0 = s0[1:] * float(h)
    assert pg[0] == 0.0
    ref_pg = -45 / (np.pi * float(h) ** 6) * (float(h) - r0) ** 2 / r0
    assert np.allclose(pg[1:300], ref_pg, rtol=1e-5, atol=1e-6 * np.abs(ref_pg).max())


# -------------------------------------------------------------- z-index
def _interleave(x, y, z):
    k = 0
    for i in range(10):
        k |= ((x >> i) & 1) << (3 * i) | ((y >> i) & 1) << (3 * i + 1) | ((z >> i) & 1) << (3 * i + 2)
    return k


def test_ztable_closed_form():
    t = np.zeros(3072, np.uint32)
    ol.oracle().orc_ztable(ptr(t, C.c_uint32))
    t = t.reshape(1024, 3)
    for k in (0, 1, 2, 3, 5, 511, 512, 1000, 1023):
        assert t[k, 0] == _interleave(k, 0, 0) and t[k, 1] == _interleave(0, k, 0) and t[k, 2] == _interleave(0, 0, k)


@needs_ref
def test_ztable_matches_reference():
    a = np.zeros(3072, np.uint32)
    b = np.zeros(3072, np.uint32)
    ol.oracle().orc_ztable(ptr(a, C.c_uint32))
    ol.ref_lib().ref_ztable(ptr(b, C.c_uint32))
    assert np.array_equal(a, b)


@needs_ref
def test_block_size_and_box_query_match_reference():
    rng = np.random.RandomState(1)
    for _ in range(3000):
        size = int(rng.randint(1, 70))
        mn = rng.randint(0, 900, 3).astype(np.uint16)
        mx = np.minimum(mn + rng.randint(0, 130, 3), 1023).astype(np.uint16)
        out = np.zeros(2 * 4096, np.uint32)
        bs = np.zeros(1, np.uint32)
        m = ol.ref_lib().ref_box_query(size, mn.ctypes.data_as(C.POINTER(C.c_ushort)), mx.ctypes.data_as(C.POINTER(C.c_ushort)),
                              ptr(out, C.c_uint32), 4096, ptr(bs, C.c_uint32))
        if m > 4096:
            continue
        ref = sorted(map(tuple, out[:2 * m].reshape(m, 2)))
        o = np.zeros(2 * 4096, np.uint32)
        k = ol.oracle().orc_box_query(int(bs[0]), ptr(mn.astype(np.uint32), C.c_uint32),
                                      ptr(mx.astype(np.uint32), C.c_uint32), ptr(o, C.c_uint32), 4096)
        mine = list(map(tuple, o[:2 * k].reshape(k, 2)))
        assert mine == ref
        # the reference's block size: smallest power of two >= size
        assert bs[0] >= size and (bs[0] == 1 or bs[0] // 2 < size)


# ---------------------------------------------------------- rigid bodies
def _body_state(rng):
    st = np.zeros(47, np.float32)
    st[0:3] = rng.uniform(0.2, 0.8, 3)
    st[9:12] = rng.normal(0, 0.05, 3)
    st[12:15] = rng.normal(0, 1e-4, 3)
    q = rng.normal(size=4)
    st[15:19] = q / np.linalg.norm(q)
    w, x, y, z = st[15:19].astype(np.float64)
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    st[19:28] = R.ravel()
    I = np.diag(rng.uniform(1e-4, 1e-3, 3))
    st[28:37] = np.linalg.inv(I).ravel()
    st[46] = rng.uniform(0.5, 3.0)
    return st


@needs_ref
def test_rigid_update_matches_reference():
    rng = np.random.RandomState(2)
    for trial in range(20):
        st = _body_state(rng)
        n = 50
        pos = (st[0:3] + rng.uniform(-0.05, 0.05, (n, 3))).astype(np.float32)
        f = rng.normal(0, 0.01, (n, 3)).astype(np.float32)
        m = np.full(n, st[46] / n, np.float32)
        a, b = st.copy(), st.copy()
        for _ in range(5):
            ol.ref_lib().ref_rigid_step(ptr(a), n, ptr(pos), ptr(f), ptr(m), 0.001, 9.81, 20.0)
            ol.oracle().orc_rigid_step(ptr(b), n, ptr(pos), ptr(f), ptr(m), 0.001, 9.81, 20.0, 0)
        assert np.array_equal(a, b), trial


def test_rigid_free_fall_closed_form():
    """No particle force: x = x0 - g t^2 / 2 (symplectic Euler: - g dt^2 k(k+1)/2)."""
    rng = np.random.RandomState(4)
    st = _body_state(rng)
    st[9:15] = 0.0
    n = 10
    pos = np.tile(st[0:3], (n, 1)).astype(np.float32)
    f = np.zeros((n, 3), np.float32)
    m = np.full(n, st[46] / n, np.float32)
    y0 = float(st[1])
    dt, g, k = 0.001, 9.81, 50
    for _ in range(k):
        ol.oracle().orc_rigid_step(ptr(st), n, ptr(pos), ptr(f), ptr(m), dt, g, 20.0, 0)
    assert abs((y0 - st[1]) - g * dt * dt * k * (k + 1) / 2) < 1e-5
    assert np.all(st[6:9] == 0.0)  # no torque, no spin


# ------------------------------------------------------ neighbours, sort
def _small_scene(n=8, seed=0):
    p, pos, typ, nf = upf.standard_dam_break(n, n, n, seed=seed)
    return p, pos, typ, nf


def test_neighbor_lists_equal_brute_force():
    p, pos, typ, nf = _small_scene(8)
    o = ol.OracleSim(p, pos, None, typ)
    st = o.state()
    counts, flat = o.neighbors()
    # brute force over the sorted state (the oracle's particle order)
    x = np.zeros((len(typ), 3), np.float32)
    x[np.arange(len(typ))] = st["pos"][st["ids"], :3]
    d2 = ((x[:, None, :] - x[None, :, :]) ** 2).astype(np.float32).sum(-1, dtype=np.float32)
    dist = np.sqrt(d2)
    off = 0
    h = np.float32(p.support_radius)
    for i in range(len(typ)):
        ref = np.flatnonzero(dist[i] < h)
        ref = ref[ref != i]
        assert np.array_equal(flat[off:off + counts[i]], ref), i
        off += counts[i]


def test_sqrt_threshold_is_exact():
    """d2 < support_radius_sq_limit  <=>  sqrtf(d2) < h, at the limit."""
    p = upf.make_params(0.01)
    T = np.float32(p.support_radius_sq_limit)
    h = np.float32(p.support_radius)
    below = np.nextafter(T, np.float32(0))
    assert np.sqrt(T) >= h and np.sqrt(below) < h


def test_oracle_stable_sort():
    rng = np.random.RandomState(5)
    k = rng.randint(0, 50, 10000).astype(np.uint32)
    out = np.zeros_like(k)
    perm = np.zeros(len(k), np.int32)
    ol.oracle().orc_stable_sort(len(k), ptr(k, C.c_uint32), ptr(out, C.c_uint32), ptr(perm, C.c_int))
    assert np.array_equal(perm, np.argsort(k, kind="stable"))


# ---------------------------------------------------- density error factor
def test_density_error_factor_product_equals_oracle():
    for dt in (0.001, 0.0012, 0.0005):
        p, pos, typ, nf = _small_scene(8)
        p.dt = dt
        q = p.copy()
        upf.density_error_factor(p)
        ol.oracle().orc_density_error_factor(C.byref(q), None)
        assert p.density_error_factor == q.density_error_factor
        assert p.density_error_factor_parameter == q.density_error_factor_parameter


def test_density_error_factor_closed_form_lattice():
    """delta = 0.05 (dt/0.0002)^2 / (dt^2 V0^2 * 2 sum_j |gradW_j|^2) for a
    particle with a full lattice neighbourhood (sum gradW = 0; the reference's
    pairwise accumulation counts every pair twice), computed in float64."""
    p, pos, typ, nf = _small_scene(8)
    upf.density_error_factor(p)
    h = float(p.support_radius)
    dx = float(p.spacing)
    offs = np.array([(i, j, k) for i in range(-3, 4) for j in range(-3, 4) for k in range(-3, 4)], np.float64) * dx
    r = np.linalg.norm(offs, axis=1)
    sel = (r > 0) & (r < h)
    g = -45 / (np.pi * h ** 6) * (h - r[sel]) ** 2 / r[sel]
    grad2 = np.sum((g[:, None] * offs[sel]) ** 2)
    V0 = float(p.initial_mass) / 1000.0
    dt = float(p.dt)
    delta = 0.05 * (dt / 0.0002) ** 2 / (dt * dt * V0 * V0 * 2 * grad2)
    assert abs(p.density_error_factor - delta) / delta < 1e-4


# ---------------------------------------------------- step invariants
def _fluid_block(n=10, lattice=1.0025):
    """A free fluid block (no walls), zero gravity, at rest density."""
    p = upf.make_params(0.01, 0.001)
    pos, typ, nf = upf.dam_break(p, n, n, n, (0.02 * n, 0.02 * n, 0.02 * n), wall_rel=1.0, lattice_rel=lattice)
    keep = typ == upf.UPF_LIQUID
    pos, typ = pos[keep], typ[keep]
    upf.density_error_factor(p)
    p.gravity = 0.0
    for a in range(3):  # no box clamping: the block floats freely
        p.collision_min[a] = 0.0
        p.collision_max[a] = p.vbox_max[a] * 0.99
    return p, pos, typ


def test_pcisph_momentum_conserved_without_walls_and_gravity():
    p, pos, typ = _fluid_block(10, lattice=0.995)  # compressed: pressure forces act
    rng = np.random.RandomState(6)
    vel = np.zeros_like(pos)
    vel[:, :3] = rng.normal(0, 0.05, (len(pos), 3))
    o = ol.OracleSim(p, pos, vel, typ, precision=64)
    P0 = vel[:, :3].astype(np.float64).sum(0)
    o.step(1, upf.UPF_PCISPH_STATIC)
    st = o.state()
    P1 = st["vel"][:, :3].astype(np.float64).sum(0)
    K, r = o.stats()
    assert K[0] >= 3 and r[0] > 1000.0
    scale = np.abs(st["force"][:, :3]).sum() * float(p.dt) / float(p.initial_mass)
    assert np.abs(P1 - P0).max() < 1e-6 * scale


def test_isolated_particle_density_is_self_contribution():
    p = upf.make_params(0.01, 0.001)
    pos, typ, nf = upf.dam_break(p, 1, 1, 1, (0.2, 0.2, 0.2), wall_rel=1.0)
    keep = typ == upf.UPF_LIQUID
    pos, typ = pos[keep], typ[keep]
    upf.density_error_factor(p)
    p.gravity = 0.0
    o = ol.OracleSim(p, pos, None, typ)
    o.step(1, upf.UPF_PCISPH_STATIC)
    st = o.state()
    assert st["rho"][0] == np.float32(np.float32(p.kernel_self) * np.float32(p.initial_mass))


def test_fp32_oracle_tracks_fp64_twin():
    """One step of config-0 layout: FP32 vs FP64 restatement (error statement)."""
    p, pos, typ, nf = _small_scene(12)
    a = ol.OracleSim(p, pos, None, typ, precision=32)
    b = ol.OracleSim(p, pos, None, typ, precision=64)
    a.step(1, upf.UPF_PCISPH_STATIC)
    b.step(1, upf.UPF_PCISPH_STATIC)
    sa, sb = a.state(), b.state()
    f = typ == upf.UPF_LIQUID
    # The FP32 and FP64 runs see the same tabulated kernel, but a distance
    # within rounding of a table bin edge falls into different bins, and the
    # first step of the dam break runs many iterations; the error statement
    # is therefore: median |dx| below 1e-7 m, max below 1e-4 m (0.01 spacing).
    dx = np.abs(sa["pos"][:, :3] - sb["pos"][:, :3])
    assert np.median(dx) < 1e-7 and dx.max() < 1e-4
    rel = np.abs(sa["rho"][f] - sb["rho"][f]) / sb["rho"][f]
    assert np.median(rel) < 1e-5 and rel.max() < 2e-2


# --------------------------------------------------- golden vectors
def test_golden_vectors_from_reference_units():
    """Vectors written by tests/golden/make_golden.py from oracle/_ref; they
    pin the oracle where the reference tree is absent."""
    path = os.path.join(GOLDEN, "reference_units.json")
    if not os.path.exists(path):
        pytest.skip("golden vectors not generated")
    g = json.load(open(path))
    for case in g["luts"]:
        mine, ks = orc_luts(np.float32(case["h"]))
        assert np.array_equal(mine.view(np.uint32), np.array(case["bits"], np.uint32))
        assert np.float32(ks).view(np.uint32) == case["kernel_self_bits"]
    for case in g["constants"]:
        p = Params()
        ol.oracle().orc_params_init(case["spacing"], C.byref(p))
        got = [p.spacing, p.particle_radius, p.support_radius, p.force_distance, p.initial_mass]
        assert np.array_equal(np.float32(got).view(np.uint32), np.array(case["bits"][:5], np.uint32))
    for case in g["rigid"]:
        st = np.array(case["state_in"], np.uint32).view(np.float32).copy()
        pos = np.array(case["pos"], np.uint32).view(np.float32).copy()
        f = np.array(case["force"], np.uint32).view(np.float32).copy()
        m = np.array(case["mass"], np.uint32).view(np.float32).copy()
        for _ in range(case["steps"]):
            ol.oracle().orc_rigid_step(ptr(st), len(m), ptr(pos), ptr(f), ptr(m), 0.001, 9.81, 20.0, 0)
        assert np.array_equal(st.view(np.uint32), np.array(case["state_out"], np.uint32))
