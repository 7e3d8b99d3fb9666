"""Error conditions the reference aborts on, reported as codes by the device
step and by the oracle alike: a neighbour list longer than its capacity and a
boundary particle with no boundary neighbour (CPU boundary mode).  Every
kernel keeps its writes inside the lists' capacity, so the step ends normally
and the code comes back from the step call."""
import numpy as np
import pytest

import oracle_lib as ol
import scenes
import unifiedparticleframeworkcuda_amd as upf

pytestmark = pytest.mark.gpu


def _oracle_code(p, pos, typ):
    o = ol.OracleSim(p, pos, None, typ)
    with pytest.raises(RuntimeError) as e:
        o.step(1, upf.UPF_PCISPH_STATIC)
    return int(str(e.value).rsplit(":", 1)[1])


@pytest.mark.parametrize("stage", [0, 32])
def test_neighbour_overflow_reported(gpu, stage):
    """GetNeighbors (src/cpu/UnifiedPhysics.cpp:2242-2295) with a capacity of
    8 entries: interior fluid particles have about 35 neighbours."""
    p, pos, typ, nf = scenes.dam_break(16)
    p.max_neighbors = 8
    sim = upf.Simulation(p, len(typ))
    t = sim.tuning()
    t.nbr_stage = stage
    sim.set_tuning(t)
    sim.upload(pos, typ)
    with pytest.raises(upf.UPFError) as e:
        sim.step(1, upf.UPF_PCISPH_STATIC)
    assert e.value.code == upf._abi.UPF_ERR_NBR_OVERFLOW
    assert _oracle_code(p, pos, typ) == upf._abi.UPF_ERR_NBR_OVERFLOW


def _with_lone_wall_particle(p, pos, typ):
    """The dam break plus one wall particle in the empty part of the box,
    farther than the support radius from every other wall particle."""
    walls = pos[typ == upf.UPF_FROZEN, :3]
    lo, hi = walls.min(0), walls.max(0)
    x = np.array([0.25 * lo[0] + 0.75 * hi[0], 0.25 * lo[1] + 0.75 * hi[1], 0.5 * (lo[2] + hi[2])], np.float32)
    d = np.sqrt(((walls - x) ** 2).sum(1)).min()
    assert d > 1.5 * p.support_radius, d
    pos2 = np.vstack([pos, np.r_[x, 0.0].astype(np.float32)[None]])
    typ2 = np.r_[typ, np.int32(upf.UPF_FROZEN)].astype(np.int32)
    return pos2, typ2


def test_isolated_wall_particle_reported(gpu):
    """CalculateWeightedVolume (src/cpu/UnifiedPhysics.cpp:2703-2746) aborts
    when a boundary particle's boundary-neighbour sum is zero."""
    p, pos, typ, nf = scenes.dam_break(16)
    pos2, typ2 = _with_lone_wall_particle(p, pos, typ)
    sim = upf.Simulation(p, len(typ2))
    with pytest.raises(upf.UPFError) as e:
        sim.upload(pos2, typ2)
        sim.step(1, upf.UPF_PCISPH_STATIC)
    assert e.value.code == upf._abi.UPF_ERR_ISOLATED
    assert _oracle_code(p, pos2, typ2) == upf._abi.UPF_ERR_ISOLATED


def test_isolated_wall_particle_allowed_with_akinci_boundaries(gpu):
    """The Akinci boundary sum includes the particle itself, so a lone wall
    particle gets psi = 0.7 / W(0) and the step runs, as in the oracle."""
    p, pos, typ, nf = scenes.dam_break(16, mode=upf.UPF_BOUNDARY_AKINCI)
    pos2, typ2 = _with_lone_wall_particle(p, pos, typ)
    sim = upf.Simulation(p, len(typ2))
    sim.upload(pos2, typ2)
    sim.step(1, upf.UPF_PCISPH_STATIC)
    o = ol.OracleSim(p, pos2, None, typ2)
    o.step(1, upf.UPF_PCISPH_STATIC)
    g, r = sim.state(), o.state()
    assert g["psi"][-1] == r["psi"][-1] and g["psi"][-1] > 0.0
    f = typ2 == upf.UPF_LIQUID
    assert np.array_equal(g["pos"][f, :3], r["pos"][f, :3])


def test_upload_after_an_error_continues(gpu):
    """A failing step leaves the handle usable: an upload clears the error
    state, and the next steps match the oracle bit for bit."""
    p, pos, typ, nf = scenes.dam_break(16)
    pos2, typ2 = _with_lone_wall_particle(p, pos, typ)
    sim = upf.Simulation(p, len(typ2))
    with pytest.raises(upf.UPFError):
        sim.upload(pos2, typ2)
        sim.step(1, upf.UPF_PCISPH_STATIC)
    sim.upload(pos, typ)
    sim.step(2, upf.UPF_PCISPH_STATIC)
    o = ol.OracleSim(p, pos, None, typ)
    o.step(2, upf.UPF_PCISPH_STATIC)
    g, r = sim.state(), o.state()
    assert np.array_equal(g["pos"][:, :3], r["pos"][:, :3])
    assert np.array_equal(g["vel"][:, :3], r["vel"][:, :3])
