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


def test_position_leaving_the_grid_on_the_last_step_is_reported(gpu):
    """A collision box reaching past the z-index grid lets a liquid particle
    leave the grid (the reference's z_table overflow); the integration of the
    step raises UPF_ERR_RANGE, and the call whose last step it was reports
    it, as the oracle does."""
    p, pos, typ, nf = scenes.dam_break(8)
    p.collision_max[1] = p.vbox_max[1] * 1.5
    # its query box inside the grid at upload; 30 m/s carries it 3 cm past
    # the top of the grid in one step
    x = np.float32([[0.5 * (p.collision_min[0] + p.collision_max[0]), p.vbox_max[1] - p.support_radius - 0.001,
                     0.5 * (p.collision_min[2] + p.collision_max[2]), 0.0]])
    pos2 = np.vstack([pos, x])
    typ2 = np.r_[typ, np.int32(upf.UPF_LIQUID)].astype(np.int32)
    vel = np.zeros_like(pos2)
    vel[-1, 1] = 30.0
    print("grid edge: particle at y %.4f, grid top %.4f" % (x[0, 1], p.vbox_max[1]), flush=True)
    sim = upf.Simulation(p, len(typ2))
    sim.upload(pos2, typ2, vel)
    with pytest.raises(upf.UPFError) as e:
        sim.step(1, upf.UPF_PCISPH_STATIC)
    assert e.value.code == upf._abi.UPF_ERR_RANGE
    o = ol.OracleSim(p, pos2, vel, typ2)
    with pytest.raises(RuntimeError) as eo:
        o.step(1, upf.UPF_PCISPH_STATIC)
    assert int(str(eo.value).rsplit(":", 1)[1]) == upf._abi.UPF_ERR_RANGE


def test_rejected_body_upload_changes_nothing(gpu):
    """Eight bodies uploaded, then an upload of two bodies whose particles
    still name bodies 2..7: UPF_ERR_ARG, and the coupled steps that follow
    run with the eight bodies exactly as a simulation that never saw the
    rejected call."""
    p, pos, typ, bid, rel, bodies, counts = upf.standard_rigid_scene(16, 16, 32, 8)
    runs = []
    for reject in (False, True):
        sim = upf.Simulation(p, len(typ))
        sim.upload(pos, typ)
        sim.upload_bodies(bodies, bid, rel)
        if reject:
            with pytest.raises(upf.UPFError) as e:
                sim.upload_bodies(bodies[:2], bid, rel)
            assert e.value.code == upf._abi.UPF_ERR_ARG
        sim.step(2, upf.UPF_PCISPH_RIGID)
        runs.append((sim.state(), sim.bodies(8)))
        sim.close()
    (a, ba), (b, bb) = runs
    assert np.array_equal(a["pos"], b["pos"]) and np.array_equal(a["vel"], b["vel"])
    for x, y in zip(ba, bb):
        assert bytes(x) == bytes(y)


def test_set_params_keeps_the_domain(gpu):
    p, pos, typ, nf = scenes.dam_break(8)
    sim = upf.Simulation(p, len(typ))
    for field, value in (("block_shift", p.block_shift + 1), ("grid_dim", p.grid_dim // 2),
                         ("block_size", p.block_size * 2)):
        q = p.copy()
        setattr(q, field, value)
        with pytest.raises(upf.UPFError) as e:
            sim.set_params(q)
        assert e.value.code == upf._abi.UPF_ERR_ARG, field
    q = p.copy()
    q.dt = 0.0008
    sim.set_params(q)
