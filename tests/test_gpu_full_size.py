"""GPU parity at the full benchmark size: the 1M dam break (128 x 64 x 128
fluid lattice, 1,048,576 fluid + 83,980 wall particles).

Two PCISPH steps on the device against two steps of the oracle (its OpenMP
build, a few seconds on the host): keys and slot order, positions,
velocities, densities, pressures and the loop statistics agree bit for bit;
and a restart file written after the two steps continues exactly.  Then 100
steps (about a minute of the OpenMP oracle), compared after 10, 25, 50 and
100 steps, with the iteration count and largest density of every step.
"""
import numpy as np
import pytest

import oracle_lib as ol
import scenes
import unifiedparticleframeworkcuda_amd as upf

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def full_scene():
    return scenes.config1_shape(1)


def test_two_steps_at_1m_bit_exact(gpu, full_scene, tmp_path):
    p, pos, typ, nf = full_scene
    assert nf == 1 << 20
    f = typ == upf.UPF_LIQUID
    sim = upf.Simulation(p, len(typ))
    sim.upload(pos, typ)
    sim.step(2, upf.UPF_PCISPH_STATIC)
    o = ol.OracleSim(p, pos, None, typ, omp=True)
    o.step(2, upf.UPF_PCISPH_STATIC)
    g, r = sim.state(), o.state()
    for k in ("pos", "vel"):
        assert np.array_equal(g[k][:, :3], r[k][:, :3]), k
    for k in ("rho", "p"):
        assert np.array_equal(g[k][f], r[k][f]), k
    kg, rg = sim.stats()
    ko, ro = o.stats()
    assert np.array_equal(kg, ko) and np.array_equal(rg, ro)
    # restart from a file after the two steps: one more step each way
    path = str(tmp_path / "restart.txt")
    sim.save_text(path, upf.UPF_TEXT_RESTART)
    b = upf.Simulation(p, len(typ))
    b.upload_restart(path)
    sim.step(1, upf.UPF_PCISPH_STATIC)
    b.step(1, upf.UPF_PCISPH_STATIC)
    pa, pb = str(tmp_path / "a.txt"), str(tmp_path / "b.txt")
    sim.save_text(pa, upf.UPF_TEXT_RESTART)
    b.save_text(pb, upf.UPF_TEXT_RESTART)
    with open(pa, "rb") as fa, open(pb, "rb") as fb:
        assert fa.read() == fb.read()


def test_hundred_steps_at_1m_bit_exact(gpu, full_scene):
    p, pos, typ, nf = full_scene
    f = typ == upf.UPF_LIQUID
    sim = upf.Simulation(p, len(typ))
    sim.upload(pos, typ)
    o = ol.OracleSim(p, pos, None, typ, omp=True)
    done = 0
    for mark in (10, 25, 50, 100):
        sim.step(mark - done, upf.UPF_PCISPH_STATIC)
        o.step(mark - done, upf.UPF_PCISPH_STATIC)
        done = mark
        g, r = sim.state(), o.state()
        kg, rg = sim.stats()
        ko, ro = o.stats()
        print("1M dam break, %d steps: iterations %d (device) / %d (oracle)" % (mark, kg.sum(), ko.sum()), flush=True)
        assert np.array_equal(kg, ko) and np.array_equal(rg, ro), mark
        for k in ("pos", "vel"):
            assert np.array_equal(g[k][:, :3], r[k][:, :3]), (mark, k)
        assert np.array_equal(g["rho"][f], r["rho"][f]), mark
    sim.close()


def test_half_spacing_walls_akinci_bit_exact(gpu):
    """The same fluid with the configuration's wall layout: walls sampled at
    half the spacing (335,400 boundary particles) and Akinci boundary volumes
    (the bench's second 1M leg) -- twelve steps bit-equal to the oracle."""
    p = upf.make_params(0.01, 0.001)
    pos, typ, nf = upf.dam_break(p, 128, 64, 128, (2.56, 0.64, 1.28), wall_rel=0.5)
    upf.density_error_factor(p)
    p.boundary_pressure_mode = upf.UPF_BOUNDARY_AKINCI
    print("half-spacing walls: %d fluid, %d boundary" % (nf, len(typ) - nf), flush=True)
    assert nf == 1 << 20 and len(typ) - nf == 335400
    f = typ == upf.UPF_LIQUID
    sim = upf.Simulation(p, len(typ))
    sim.upload(pos, typ)
    sim.step(12, upf.UPF_PCISPH_STATIC)
    o = ol.OracleSim(p, pos, None, typ, omp=True)
    o.step(12, upf.UPF_PCISPH_STATIC)
    g, r = sim.state(), o.state()
    kg, rg = sim.stats()
    ko, ro = o.stats()
    print("half-spacing walls, 12 steps: iterations %s" % [int(k) for k in kg], flush=True)
    assert np.array_equal(kg, ko) and np.array_equal(rg, ro)
    for k in ("pos", "vel"):
        assert np.array_equal(g[k][:, :3], r[k][:, :3]), k
    assert np.array_equal(g["rho"][f], r["rho"][f])
    sim.close()
