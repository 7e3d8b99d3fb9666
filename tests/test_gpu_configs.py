"""GPU parity at the workloads of the coupling, batched and WCSPH benchmarks.

* Rigid-fluid coupling (CpuParticlePhysicsVersatileCouplingPCISPH,
  src/cpu/UnifiedPhysics.cpp:4991): 524,288 fluid particles (64 x 128 x 64
  block) with 64 seeded boxes and spheres, Akinci boundary volumes, surface
  tension 0.05, adhesion 0.01 -- three steps against the OpenMP oracle.
* Batched uneven scenes: 256 dam breaks of 2,048 .. 65,536 fluid particles in
  one batch, every step one graph replay -- each scene equals its plain
  simulation bit for bit, and eight sampled scenes equal the oracle.
* WCSPH (CpuParticlePhysicsStaticBoundariesSPH with physicsType 'w',
  src/cpu/UnifiedPhysics.cpp:4180) on the 1M dam break at the CFL step
  0.253 ms, stiffness 1000, gamma 7 -- one step within 1e-5 relative, three
  within twice the FP32 oracle's distance from its FP64 twin.
"""
import numpy as np
import pytest

import oracle_lib as ol
import scenes
import unifiedparticleframeworkcuda_amd as upf

pytestmark = pytest.mark.gpu


def _body_arrays(bs):
    return {k: np.array([list(getattr(b, k)) for b in bs], np.float32)
            for k in ("pos", "vel", "angular_velocity", "linear_momentum", "angular_momentum", "quaternion", "rotation")}


def _rel_gap(a, b):
    return float(np.abs(a - b).max() / (np.abs(b).max() + 1e-30))


@pytest.mark.parametrize("mode", [upf.UPF_BOUNDARY_AKINCI, upf.UPF_BOUNDARY_CPU_REF])
def test_rigid_coupling_workload_three_steps(gpu, mode):
    p, pos, typ, bid, rel, bodies, counts = upf.standard_rigid_scene(64, 128, 64, n_bodies=64, mode=mode)
    nf, nr, nw = counts
    print("coupling workload: %d fluid, %d rigid, %d wall, 64 bodies, mode %d" % (nf, nr, nw, mode), flush=True)
    assert nf == 524288 and len(bodies) == 64
    sim = upf.Simulation(p, len(typ))
    sim.upload(pos, typ)
    sim.upload_bodies(bodies, bid, rel)
    sim.step(3, upf.UPF_PCISPH_RIGID)
    o = ol.OracleSim(p, pos, None, typ, omp=True)
    o.upload_bodies(bodies, bid, rel, rel[:, 3].copy())
    o.step(3, upf.UPF_PCISPH_RIGID)
    g, r = sim.state(), o.state()
    # positions: relative to the container's extent; velocities, forces: to their largest magnitude
    ext = float(np.max(np.abs(r["pos"][:, :3])))
    assert np.abs(g["pos"][:, :3] - r["pos"][:, :3]).max() <= 1e-5 * ext
    assert _rel_gap(g["vel"][:, :3], r["vel"][:, :3]) <= 1e-5
    f = typ == upf.UPF_LIQUID
    assert _rel_gap(g["rho"][f], r["rho"][f]) <= 1e-5
    assert _rel_gap(g["p"][f], r["p"][f]) <= 1e-5
    gb, rb = _body_arrays(sim.bodies(64)), _body_arrays(o.bodies(64))
    for k in gb:
        scale = np.abs(rb[k]).max() + 1e-12
        assert np.abs(gb[k] - rb[k]).max() <= 1e-5 * scale, k
    assert np.array_equal(sim.stats()[0], o.stats()[0])
    bit_exact = np.array_equal(g["pos"][:, :3], r["pos"][:, :3]) and np.array_equal(g["vel"][:, :3], r["vel"][:, :3])
    print("coupling workload: K %s, bit-exact %s" % (list(sim.stats()[0]), bit_exact), flush=True)


def _split(counts, arr):
    offs = np.cumsum([0] + list(counts))
    return [arr[offs[k]:offs[k + 1]] for k in range(len(counts))]


def test_batched_workload_graph_equals_plain_and_oracle(gpu):
    nsteps = 2
    sc = upf.batch_dam_breaks(256, seed=0)
    counts = [len(s[2]) for s in sc]
    nf = sum(s[3] for s in sc)
    print("batched workload: 256 scenes, %d fluid, %d particles" % (nf, sum(counts)), flush=True)
    b = upf.Batch.from_scenes(sc)
    b.step(nsteps, use_graph=True)
    st = b.state()
    K, rmax = b.stats()
    b.close()
    assert K.shape == (nsteps, 256)
    pos_s, vel_s, rho_s = _split(counts, st["pos"]), _split(counts, st["vel"]), _split(counts, st["rho"])
    for k, (p, pos, typ, nfk) in enumerate(sc):
        sim = upf.Simulation(p, len(typ))
        sim.upload(pos, typ)
        sim.step(nsteps, upf.UPF_PCISPH_STATIC)
        g = sim.state()
        Kp, rp = sim.stats()
        sim.close()
        f = typ == upf.UPF_LIQUID
        assert np.array_equal(pos_s[k][:, :3], g["pos"][:, :3]), "scene %d positions" % k
        assert np.array_equal(vel_s[k][:, :3], g["vel"][:, :3]), "scene %d velocities" % k
        assert np.array_equal(rho_s[k][f], g["rho"][f]), "scene %d densities" % k
        assert list(K[:, k]) == list(Kp) and np.array_equal(rmax[:, k], rp), "scene %d loop statistics" % k
    rng = np.random.RandomState(7)
    for k in sorted(rng.choice(256, 8, replace=False)):
        p, pos, typ, nfk = sc[k]
        o = ol.OracleSim(p, pos, None, typ, omp=True)
        o.step(nsteps, upf.UPF_PCISPH_STATIC)
        r = o.state()
        f = typ == upf.UPF_LIQUID
        assert np.array_equal(pos_s[k][:, :3], r["pos"][:, :3]), "scene %d positions vs oracle" % k
        assert np.array_equal(vel_s[k][:, :3], r["vel"][:, :3]), "scene %d velocities vs oracle" % k
        assert np.array_equal(rho_s[k][f], r["rho"][f]), "scene %d densities vs oracle" % k
        assert list(K[:, k]) == list(o.stats()[0]), "scene %d iteration counts vs oracle" % k


def test_wcsph_1m_three_steps(gpu):
    """One step: densities, pressures, forces and positions within 1e-5
    relative (the density pass is bit-exact; powf may round the pressure
    differently in the last place).  Two more steps: positions within 1e-5 of
    the container, and velocities, densities, pressures and forces no farther
    from the FP32 oracle than twice the FP32 oracle's own distance from its
    FP64 twin -- a last-place difference can move a pair across a kernel
    table bin, which the following steps carry on, like any rounding."""
    p, pos, typ, nf = scenes.config1_shape(1)
    p.dt = 0.000253
    p.gas_const_wcsph = 1000.0
    print("WCSPH 1M: %d fluid, %d particles, dt %g" % (nf, len(typ), p.dt), flush=True)
    f = typ == upf.UPF_LIQUID
    sim = upf.Simulation(p, len(typ))
    sim.upload(pos, typ)
    o = ol.OracleSim(p, pos, None, typ, omp=True)
    o64 = ol.OracleSim(p, pos, None, typ, precision=64, omp=True)
    sim.step(1, upf.UPF_WCSPH_STATIC)
    o.step(1, upf.UPF_WCSPH_STATIC)
    g, r = sim.state(), o.state(pcisph=False)
    ext = float(np.max(np.abs(r["pos"][:, :3])))
    assert np.abs(g["pos"][:, :3] - r["pos"][:, :3]).max() <= 1e-5 * ext
    assert np.array_equal(g["rho"][f], r["rho"][f])
    for k in ("p",):
        assert _rel_gap(g[k][f], r[k][f]) <= 1e-5, k
    assert _rel_gap(g["force"][f, :3], r["force"][f, :3]) <= 1e-5
    o64.step(1, upf.UPF_WCSPH_STATIC)
    sim.step(2, upf.UPF_WCSPH_STATIC)
    o.step(2, upf.UPF_WCSPH_STATIC)
    o64.step(2, upf.UPF_WCSPH_STATIC)
    g, r, d = sim.state(), o.state(pcisph=False), o64.state(pcisph=False)
    assert np.abs(g["pos"][:, :3] - r["pos"][:, :3]).max() <= 1e-5 * ext
    for k, sl in (("vel", (f, slice(0, 3))), ("rho", f), ("p", f), ("force", (f, slice(0, 3)))):
        gap, spread = _rel_gap(g[k][sl], r[k][sl]), _rel_gap(r[k][sl], d[k][sl])
        print("WCSPH 1M after 3 steps: %s device-oracle %.2e, FP32-FP64 spread %.2e" % (k, gap, spread), flush=True)
        assert gap <= max(2 * spread, 1e-6), k
