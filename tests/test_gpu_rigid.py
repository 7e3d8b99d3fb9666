"""GPU parity of the rigid-fluid coupling step (CpuParticlePhysicsVersatileCouplingPCISPH,
src/cpu/UnifiedPhysics.cpp:4991) against the oracle.

Body forces and torques are summed in slot order, the order of the
reference's rigid_particle_indices_ (one block per body loads the members'
terms in parallel; each of the six sums then runs member by member), so the
state agrees bit for bit as long as the device's cosf / sinf round like the
host's; the contact test therefore allows 1e-5 relative.  The bodies hold
862 to 4706 members, so the reduction runs over many 256-member chunks."""
import numpy as np
import pytest

import unifiedparticleframeworkcuda_amd as upf
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _scene(mode, drop=0.0, nb=4):
    p, pos, typ, bid, rel, bodies, counts = upf.standard_rigid_scene(16, 16, 32, nb, mode=mode)
    if drop:
        r = typ == upf.UPF_RIGID
        pos[r, 1] -= drop
        for b in bodies:
            b.pos[1] -= drop
    return p, pos, typ, bid, rel, bodies


def _run(mode, steps, drop=0.0, nb=4):
    p, pos, typ, bid, rel, bodies = _scene(mode, drop, nb)
    sim = upf.Simulation(p, len(typ))
    sim.upload(pos, typ)
    sim.upload_bodies(bodies, bid, rel)
    o = ol.OracleSim(p, pos, None, typ)
    o.upload_bodies(bodies, bid, rel, rel[:, 3].copy())
    sim.step(steps, upf.UPF_PCISPH_RIGID)
    o.step(steps, upf.UPF_PCISPH_RIGID)
    return sim, o, typ, nb


def _body_arrays(bs):
    return {k: np.array([list(getattr(b, k)) for b in bs], np.float32)
            for k in ("pos", "vel", "angular_velocity", "linear_momentum", "angular_momentum", "quaternion", "rotation")}


@pytest.mark.parametrize("mode", [upf.UPF_BOUNDARY_CPU_REF, upf.UPF_BOUNDARY_AKINCI])
def test_rigid_free_fall_steps_bit_exact(gpu, mode):
    sim, o, typ, nb = _run(mode, 5)
    g, r = sim.state(), o.state()
    assert np.array_equal(g["pos"][:, :3], r["pos"][:, :3])
    assert np.array_equal(g["vel"][:, :3], r["vel"][:, :3])
    f = typ != upf.UPF_FROZEN
    assert np.array_equal(g["force"][f, :3], r["force"][f, :3])
    gb, rb = _body_arrays(sim.bodies(nb)), _body_arrays(o.bodies(nb))
    for k in gb:
        assert np.array_equal(gb[k], rb[k]), k
    assert np.array_equal(sim.stats()[0], o.stats()[0])


@pytest.mark.parametrize("mode", [upf.UPF_BOUNDARY_CPU_REF, upf.UPF_BOUNDARY_AKINCI])
def test_rigid_contact_steps(gpu, mode):
    """Bodies lowered onto the fluid: coupling forces, torques and rotation."""
    sim, o, typ, nb = _run(mode, 8, drop=0.12)
    g, r = sim.state(), o.state()
    dx = np.abs(g["pos"][:, :3] - r["pos"][:, :3]).max()
    assert dx < 1e-5 * 0.5, dx
    gb, rb = _body_arrays(sim.bodies(nb)), _body_arrays(o.bodies(nb))
    assert np.abs(rb["angular_momentum"]).max() > 0  # the bodies do spin
    for k in gb:
        scale = np.abs(rb[k]).max() + 1e-12
        assert np.abs(gb[k] - rb[k]).max() <= 1e-5 * scale, k


@pytest.mark.parametrize("mode", [upf.UPF_BOUNDARY_CPU_REF, upf.UPF_BOUNDARY_AKINCI])
def test_bodies_built_from_particles_drive_the_coupling_step(gpu, mode):
    """The scene's rigid particles solidified into bodies as the reference's
    scene set-up does (upf_bodies_from_particles: AddToRigidBody's centre of
    mass, inertia and momenta, RigidBody()'s zero momentum and world inverse
    inertia): 5 coupled steps agree with the oracle bit for bit."""
    p, pos, typ, bid0, rel0, bodies0 = _scene(mode, nb=4)
    r = typ == upf.UPF_RIGID
    b_r, rel_r, bodies = upf.bodies_from_particles(p, pos[r], rel0[r, 3])
    assert len(bodies) == len(bodies0) and np.all(b_r >= 0)
    for b, b0 in zip(bodies, bodies0):
        assert b.mass == pytest.approx(b0.mass, rel=1e-4)
        assert np.allclose(list(b.pos), list(b0.pos), atol=1e-3)
    bid = np.full(len(typ), -1, np.int32)
    bid[r] = b_r
    rel = np.zeros((len(typ), 4), np.float32)
    rel[r] = rel_r
    sim = upf.Simulation(p, len(typ))
    sim.upload(pos, typ)
    sim.upload_bodies(bodies, bid, rel)
    o = ol.OracleSim(p, pos, None, typ)
    o.upload_bodies(bodies, bid, rel, rel[:, 3].copy())
    sim.step(5, upf.UPF_PCISPH_RIGID)
    o.step(5, upf.UPF_PCISPH_RIGID)
    g, rr = sim.state(), o.state()
    assert np.array_equal(g["pos"][:, :3], rr["pos"][:, :3])
    assert np.array_equal(g["vel"][:, :3], rr["vel"][:, :3])
    gb, rb = _body_arrays(sim.bodies(len(bodies))), _body_arrays(o.bodies(len(bodies)))
    for k in gb:
        assert np.array_equal(gb[k], rb[k]), k
