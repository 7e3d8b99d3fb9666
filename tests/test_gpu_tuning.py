"""Launch tuning changes speed only: every tuning gives bit-identical steps
(including list-build staging overflow), invalid shapes are rejected, and the
autotuner runs, applies and caches its choice."""
import ctypes as C

import numpy as np
import pytest

import unifiedparticleframeworkcuda_amd as upf
from unifiedparticleframeworkcuda_amd import autotune
from unifiedparticleframeworkcuda_amd._abi import Tuning

pytestmark = pytest.mark.gpu


def _compressed_scene():
    # lattice at 0.9 spacing: ~45 neighbours per particle, past a 32-entry stage
    p = upf.make_params(0.01, 0.001)
    pos, typ, nf = upf.dam_break(p, 14, 10, 12, (0.28, 0.2, 0.2), lattice_rel=0.9)
    upf.density_error_factor(p)
    return p, pos, typ


def _run(p, pos, typ, tuning, steps=5):
    sim = upf.Simulation(p, len(typ))
    sim.set_tuning(tuning)
    sim.upload(pos, typ)
    sim.step(steps, upf.UPF_PCISPH_STATIC)
    st = sim.state()
    K, _ = sim.stats()
    sim.close()
    return st, K


@pytest.mark.parametrize("tuning", [
    dict(loop_block=64, grid_cap=16, nbr_rounds=2, nbr_stage=32, speculate=1, loop_list=0, loop_width=4, nbr_tile=0,
         spec_first=0, xcd_remap=0, nbr_skip=1, nbr_fuse=1),
    dict(loop_block=128, grid_cap=2048, nbr_rounds=8, nbr_stage=32, speculate=6, loop_list=1, loop_width=8,
         nbr_tile=256, spec_first=64, xcd_remap=1, nbr_skip=0, nbr_fuse=1),
    dict(loop_block=256, grid_cap=3, nbr_rounds=4, nbr_stage=64, speculate=2, loop_list=0, loop_width=8, nbr_tile=512,
         spec_first=2, xcd_remap=1, nbr_skip=1, nbr_fuse=0),
    dict(loop_block=256, grid_cap=8192, nbr_rounds=8, nbr_stage=0, speculate=4, loop_list=1, loop_width=4, nbr_tile=0,
         spec_first=4, xcd_remap=0, nbr_skip=1, nbr_fuse=0),
    # union lists: neighbour records from LDS (chunks whose union is too large
    # for it fall back to the delta-coded lists); a small grid re-stages per chunk
    dict(loop_block=256, grid_cap=8192, nbr_rounds=4, nbr_stage=0, speculate=4, loop_list=2, loop_width=8, nbr_tile=0,
         spec_first=4, xcd_remap=1, nbr_skip=0, nbr_fuse=1),
    dict(loop_block=256, grid_cap=3, nbr_rounds=4, nbr_stage=0, speculate=1, loop_list=2, loop_width=4, nbr_tile=0,
         spec_first=0, xcd_remap=0, nbr_skip=0, nbr_fuse=1),
])
def test_tunings_give_identical_steps(gpu, tuning):
    p, pos, typ = _compressed_scene()
    print("tuning %s, %d particles" % (tuning, len(typ)), flush=True)
    ref, Kr = _run(p, pos, typ, autotune.DEFAULT)
    got, Kg = _run(p, pos, typ, tuning)
    assert list(Kr) == list(Kg)
    for k in ("pos", "vel", "rho", "p", "force"):
        assert np.array_equal(ref[k], got[k]), k


def _wide_cell_scene():
    """A dam break whose virtual box is chosen so that cells come out 1.8 h
    wide (scaled h just above a power of two: the block doubles)."""
    p, pos, typ, nf = upf.standard_dam_break(20, 14, 16)
    h = p.support_radius
    e_min = float(p.vbox_max[0])
    k = 10
    while 1024.0 * h / (2 ** k * 1.1) < e_min:
        k -= 1
    e = 1024.0 * h / (2 ** k * 1.1)
    vmax = np.array([e, e, e], np.float32)
    cmin = np.array(list(p.collision_min), np.float32)
    cmax = np.array(list(p.collision_max), np.float32)
    upf.check("upf_params_set_domain", upf.lib().upf_params_set_domain(
        C.byref(p), vmax.ctypes.data_as(C.POINTER(C.c_float)), cmin.ctypes.data_as(C.POINTER(C.c_float)),
        cmax.ctypes.data_as(C.POINTER(C.c_float))))
    upf.density_error_factor(p)
    return p, pos, typ, nf


def test_round_skip_on_wide_cells_equals_the_oracle(gpu):
    """A scene whose cells are much wider than h (cell side 1.8 h: about 47
    particles per cell, most rounds of 64 candidates miss a particle's query
    box): the list build that passes over such rounds gives the steps of the
    one that tests every round, and both equal the oracle."""
    import oracle_lib as ol
    p, pos, typ, nf = _wide_cell_scene()
    ratio = (1 << p.block_shift) / p.scales[0] / p.support_radius
    print("wide cells: %d particles, cell side %.2f h" % (len(typ), ratio), flush=True)
    assert ratio > 1.5
    runs = [_run(p, pos, typ, dict(autotune.DEFAULT, nbr_rounds=r, nbr_skip=k, nbr_fuse=f), steps=3)
            for r in (2, 8) for k in (0, 1) for f in (0, 1)]
    o = ol.OracleSim(p, pos, None, typ)
    o.step(3, upf.UPF_PCISPH_STATIC)
    r = o.state()
    for got, K in runs:
        assert list(K) == list(o.stats()[0])
        for k in ("pos", "vel"):
            assert np.array_equal(got[k][:, :3], r[k][:, :3]), k


def test_invalid_tuning_rejected(gpu):
    p, pos, typ = _compressed_scene()
    sim = upf.Simulation(p, len(typ))
    for d in (dict(autotune.DEFAULT, loop_block=96), dict(autotune.DEFAULT, nbr_rounds=3),
              dict(autotune.DEFAULT, nbr_stage=16), dict(autotune.DEFAULT, grid_cap=0),
              dict(autotune.DEFAULT, loop_list=3), dict(autotune.DEFAULT, loop_list=2, loop_block=128),
              dict(autotune.DEFAULT, loop_width=6),
              dict(autotune.DEFAULT, nbr_tile=128), dict(autotune.DEFAULT, nbr_skip=2),
              dict(autotune.DEFAULT, nbr_fuse=2)):
        with pytest.raises(upf.UPFError):
            sim.set_tuning(d)
    assert sim.tuning().as_dict() == autotune.DEFAULT
    sim.close()


def test_autotuner_applies_and_caches(gpu, tmp_path):
    p, pos, typ, nf = upf.standard_dam_break(16, 12, 16)
    sim = upf.Simulation(p, len(typ))
    path = str(tmp_path / "tuning.json")
    best, ms = autotune.tune_simulation(sim, pos, typ, "dam_break_test", steps=2, warmup=1, path=path)
    assert sim.tuning().as_dict() == best and ms > 0
    assert autotune.load("dam_break_test", len(typ), path) == best
    again, ms2 = autotune.tune_simulation(sim, pos, typ, "dam_break_test", path=path)
    assert again == best and ms2 is None  # from the cache
    sim.close()


@pytest.mark.parametrize("mode", [upf.UPF_BOUNDARY_CPU_REF, upf.UPF_BOUNDARY_AKINCI])
@pytest.mark.parametrize("lattice_rel", [1.0025, 0.9])
def test_union_lists_equal_delta_coded_lists(gpu, mode, lattice_rel):
    """Union lists (loop_list 2) against delta-coded lists (loop_list 1) on the
    dam break at rest spacing (unions fit in LDS) and compressed (some do
    not): the same steps bit for bit, in both boundary-pressure modes."""
    p = upf.make_params(0.01, 0.001)
    pos, typ, nf = upf.dam_break(p, 24, 16, 20, (0.48, 0.32, 0.4), lattice_rel=lattice_rel)
    upf.density_error_factor(p)
    p.boundary_pressure_mode = mode
    ref, Kr = _run(p, pos, typ, dict(autotune.DEFAULT, loop_list=1), steps=6)
    got, Kg = _run(p, pos, typ, dict(autotune.DEFAULT, loop_list=2), steps=6)
    assert list(Kr) == list(Kg)
    for k in ("pos", "vel", "rho", "p", "force"):
        assert np.array_equal(ref[k], got[k]), k


def test_fused_list_build_for_wcsph_and_coupling_equals_separate_kernels(gpu):
    """The list build fused with the WCSPH / original SPH density pass, and
    with the rigid-fluid coupling pass (both boundary modes), steps exactly
    like the separate kernels."""
    def run(method, make, fuse, steps):
        p, pos, typ, extra = make()
        sim = upf.Simulation(p, len(typ))
        sim.set_tuning(dict(autotune.DEFAULT, nbr_fuse=fuse))
        sim.upload(pos, typ)
        if extra is not None:
            sim.upload_bodies(*extra)
        sim.step(steps, method)
        st = sim.state()
        sim.close()
        return st

    def wcsph():
        p, pos, typ, nf = upf.standard_dam_break(16, 12, 16)
        p.dt = 0.000253
        return p, pos, typ, None

    def rigid(mode):
        def make():
            p, pos, typ, bid, rel, bodies, counts = upf.standard_rigid_scene(16, 24, 16, n_bodies=4, mode=mode)
            return p, pos, typ, (bodies, bid, rel)
        return make

    for method, make, steps in ((upf.UPF_WCSPH_STATIC, wcsph, 6), (upf.UPF_SPH_STATIC, wcsph, 4),
                                (upf.UPF_PCISPH_RIGID, rigid(upf.UPF_BOUNDARY_AKINCI), 4),
                                (upf.UPF_PCISPH_RIGID, rigid(upf.UPF_BOUNDARY_CPU_REF), 4)):
        a, b = run(method, make, 1, steps), run(method, make, 0, steps)
        print("method %d: fused vs separate kernels" % method, flush=True)
        for k in ("pos", "vel", "rho"):
            assert np.array_equal(a[k], b[k]), (method, k)


def test_rigid_lists_by_the_wave_build_equal_the_per_particle_build(gpu):
    """The rigid slots' lists of the coupling step come from the wave-cooperative
    build (nbr_tile 0) or from one thread per rigid slot over its query box
    (nbr_tile 256): both boundary modes step exactly alike.  A rigid particle's
    sort key is that of its position before the last rigid update; with the
    bodies thrown sideways at 3 m/s (3 mm per step, a seventh of a cell) into
    the container's walls, rigid particles' query boxes reach cells two away
    from their key's, where wall particles lie within h, so the build must
    centre their stencils on their current cells."""
    for mode in (upf.UPF_BOUNDARY_AKINCI, upf.UPF_BOUNDARY_CPU_REF):
        out = []
        for tile in (0, 256):
            p, pos, typ, bid, rel, bodies, counts = upf.standard_rigid_scene(16, 24, 16, n_bodies=4, mode=mode)
            for k, b in enumerate(bodies):
                b.vel[0] = 3.0 if k % 2 == 0 else -3.0
                b.linear_momentum[0] = b.vel[0] * b.mass
            sim = upf.Simulation(p, len(typ))
            sim.set_tuning(dict(autotune.DEFAULT, nbr_tile=tile))
            sim.upload(pos, typ)
            sim.upload_bodies(bodies, bid, rel)
            sim.step(60, upf.UPF_PCISPH_RIGID)
            out.append(sim.state())
            sim.close()
        print("mode %d: rigid lists, wave build vs per-particle build" % mode, flush=True)
        for k in ("pos", "vel", "rho", "p", "force"):
            assert np.array_equal(out[0][k], out[1][k]), (mode, k)
