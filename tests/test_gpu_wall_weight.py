"""GPU: the wall weight function (add_wall_weight) against the oracle.

PCISPH adds it to the predicted density before the neighbour sum at the
predicted position (the GPU path's ComputePredictedDensityAndPressureBlocks-
KernelPCISPHWallWeight), WCSPH after the neighbour sum at the position (the
CPU path's CalculateDensityPressure); both must agree with the oracle as the
steps without it do (bit for bit for PCISPH, powf's last ulp for WCSPH).
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as ol
import scenes
import unifiedparticleframeworkcuda_amd as upf
from test_gpu_pipeline import _close, _compare
from test_wall_weight_oracle import _distinguishing_heights, _floor_points
from unifiedparticleframeworkcuda_amd._abi import UPF_ERR_ARG

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [upf.UPF_BOUNDARY_CPU_REF, upf.UPF_BOUNDARY_AKINCI])
def test_pcisph_with_wall_weight_bit_exact(gpu, mode):
    p, pos, typ, nf = scenes.dam_break(16, mode=mode)
    p.add_wall_weight = 1
    f = typ == upf.UPF_LIQUID
    sim = upf.Simulation(p, len(typ))
    sim.upload(pos, typ)
    sim.step(3, upf.UPF_PCISPH_STATIC)
    o = ol.OracleSim(p, pos, None, typ)
    o.step(3, upf.UPF_PCISPH_STATIC)
    _compare(sim, o, f)
    kg, rg = sim.stats()
    ko, ro = o.stats()
    assert np.array_equal(kg, ko) and np.array_equal(rg, ro)
    # the term changes the step
    q = p.copy()
    q.add_wall_weight = 0
    b = upf.Simulation(q, len(typ))
    b.upload(pos, typ)
    b.step(3, upf.UPF_PCISPH_STATIC)
    assert not np.array_equal(b.state()["rho"][f], sim.state()["rho"][f])


def test_wcsph_with_wall_weight(gpu):
    p, pos, typ, nf = scenes.dam_break(16)
    p.add_wall_weight = 1
    p.gas_const_wcsph = 1000.0
    p.dt = 0.000253
    f = typ == upf.UPF_LIQUID
    sim = upf.Simulation(p, len(typ))
    sim.upload(pos, typ)
    o = ol.OracleSim(p, pos, None, typ)
    sim.step(1, upf.UPF_WCSPH_STATIC)
    o.step(1, upf.UPF_WCSPH_STATIC)
    g, r = sim.state(), o.state(pcisph=False)
    assert np.array_equal(g["rho"][f], r["rho"][f])  # the density of the first step is exact
    sim.step(2, upf.UPF_WCSPH_STATIC)
    o.step(2, upf.UPF_WCSPH_STATIC)
    g, r = sim.state(), o.state(pcisph=False)
    # device and host powf may differ in the last ulp; the wall term raises
    # the pressure near the walls, which carries such differences into the
    # velocities: compared to within 2e-5 of the largest speed
    vmax = float(np.abs(r["vel"][f, :3]).max())
    _close(g["pos"][:, :3], r["pos"][:, :3], "pos", rtol=1e-6)
    _close(g["vel"][f, :3], r["vel"][f, :3], "vel", rtol=1e-4, atol=2e-5 * vmax)
    _close(g["rho"][f], r["rho"][f], "rho", rtol=1e-6)


@pytest.mark.parametrize("method", [upf.UPF_PCISPH_STATIC, upf.UPF_WCSPH_STATIC])
def test_lone_particles_where_the_table_index_orders_differ(gpu, method):
    """Particles too far apart to be neighbours, at rest at heights where the
    CPU path's and the GPU path's table index give different weights: the
    density of each is m W(0) + the weight of the path the method follows."""
    p, _, _, _ = scenes.dam_break(16)
    p.add_wall_weight = 1
    p.gravity = 0.0
    p.gas_const_wcsph = 1000.0
    ys = _distinguishing_heights(p, 8)
    assert len(ys) > 0
    x = _floor_points(p, ys)
    lo, hi = p.container_min[0], p.container_max[0]
    x[:, 0] = np.linspace(lo + 0.1 * (hi - lo), lo + 0.9 * (hi - lo), len(ys)).astype(np.float32)
    t = np.full(len(ys), upf.UPF_LIQUID, np.int32)
    sim = upf.Simulation(p, len(t))
    sim.upload(x, t)
    sim.step(1, method)
    o = ol.OracleSim(p, x, None, t)
    o.step(1, method)
    pcisph = method == upf.UPF_PCISPH_STATIC
    g, r = sim.state(), o.state(pcisph=pcisph)
    assert np.array_equal(g["rho"], r["rho"]) and np.array_equal(g["pos"], r["pos"])
    # WCSPH: the density of the positions the step starts from; PCISPH: the
    # predicted density at the predicted positions (moved into the collision
    # box as the step's prediction does), where the step leaves the particles
    _, w_cpu, w_gpu = ol.wall_weight(p, r["pos"] if pcisph else x)
    m = np.float32(p.initial_mass)
    w = w_gpu if pcisph else w_cpu
    assert np.array_equal(g["rho"], (np.float32(np.float32(p.kernel_self) * m) + w).astype(np.float32))


def test_wall_weight_rejected_where_unsupported(gpu):
    p, pos, typ, nf = scenes.dam_break(8)
    p.add_wall_weight = 1
    L = upf.lib()
    h = C.c_void_p()
    counts = np.array([len(typ)], np.int32)
    params = (upf.Params * 1)(p)
    assert L.upf_batch_create(1, params, upf.ptr(counts, C.c_int), 64, C.byref(h)) == UPF_ERR_ARG
    p.add_wall_weight = 2
    with pytest.raises(upf.UPFError):
        upf.Simulation(p, len(typ))
