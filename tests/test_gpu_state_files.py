"""GPU: particle files written from a running simulation (upf_save_text)
and restart from them (upf_read_restart + upf_upload_restart,
Scene::readFluidObjectFromFile, src/sceneConfigure/Scene.cpp:648-697).

The files must equal, byte for byte, the files the oracle's iostream writers
(pinned against the reference's UnifiedIO.cpp) give for the oracle's state
after the same steps; a simulation restarted from a restart file must
continue exactly as the one that wrote it.

Particle order: the device sorts at the start of a step (the reference's GPU
path, whose CopyParticlesCUDA feeds the writers), the oracle at the end (the
CPU path's particles_).  After step k the device's order is therefore the
oracle's order after step k - 1, which the tests take from the oracle.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as ol
import scenes
import unifiedparticleframeworkcuda_amd as upf
from unifiedparticleframeworkcuda_amd import UPF_TEXT_LIQUID, UPF_TEXT_RIGID, UPF_TEXT_POSITIONS, UPF_TEXT_RESTART
from unifiedparticleframeworkcuda_amd._abi import UPF_ERR_ARG, UPF_ERR_STATE

pytestmark = pytest.mark.gpu

ALL = (UPF_TEXT_LIQUID, UPF_TEXT_RIGID, UPF_TEXT_POSITIONS, UPF_TEXT_RESTART)


def read_bytes(path):
    with open(path, "rb") as f:
        return f.read()


def oracle_steps(o, k, method):
    """k steps of the oracle; returns the upload indices in the device's
    particle order after them (the oracle's order before its last sort)."""
    if k > 1:
        o.step(k - 1, method)
    ids = o.state()["ids"].copy()
    o.step(1, method)
    return ids


def oracle_files(o, ids, tmp_path, tag, kinds=ALL):
    st = o.slot_state(ids=ids)
    out = {}
    for which in kinds:
        path = str(tmp_path / ("o_%s_%d.txt" % (tag, which)))
        ol.write_text(path, which, st["pos"], st["type"], st["vel"], st["p"], st["rho"])
        out[which] = read_bytes(path)
    return out


def sim_files(sim, tmp_path, tag, kinds=ALL):
    out = {}
    for which in kinds:
        path = str(tmp_path / ("g_%s_%d.txt" % (tag, which)))
        sim.save_text(path, which)
        out[which] = read_bytes(path)
    return out


@pytest.mark.parametrize("mode", [upf.UPF_BOUNDARY_CPU_REF, upf.UPF_BOUNDARY_AKINCI])
def test_saved_files_equal_oracle_files(gpu, tmp_path, mode):
    p, pos, typ, nf = scenes.dam_break(16, mode=mode)
    sim = upf.Simulation(p, len(typ))
    sim.upload(pos, typ)
    sim.step(3, upf.UPF_PCISPH_STATIC)
    o = ol.OracleSim(p, pos, None, typ)
    ids = oracle_steps(o, 3, upf.UPF_PCISPH_STATIC)
    g = sim_files(sim, tmp_path, "a")
    r = oracle_files(o, ids, tmp_path, "a")
    for which in ALL:
        assert g[which] == r[which], which
    assert int(g[UPF_TEXT_LIQUID].split(b"\n")[0]) == nf
    assert int(g[UPF_TEXT_RIGID].split(b"\n")[0]) == len(typ) - nf


@pytest.mark.parametrize("method,mode", [(upf.UPF_PCISPH_STATIC, upf.UPF_BOUNDARY_CPU_REF),
                                         (upf.UPF_PCISPH_STATIC, upf.UPF_BOUNDARY_AKINCI),
                                         (upf.UPF_WCSPH_STATIC, upf.UPF_BOUNDARY_CPU_REF)])
def test_restart_continues_exactly(gpu, tmp_path, method, mode):
    """Run A: 3 steps, restart file, 4 more steps.  Run B: a new simulation
    loaded from that file, 4 steps.  Both end in the same bytes."""
    p, pos, typ, nf = scenes.dam_break(16, mode=mode)
    if method == upf.UPF_WCSPH_STATIC:
        p.gas_const_wcsph = 1000.0
        p.dt = 0.000253
    a = upf.Simulation(p, len(typ))
    a.upload(pos, typ)
    a.step(3, method)
    path = str(tmp_path / "restart.txt")
    a.save_text(path, UPF_TEXT_RESTART)
    a.step(4, method)
    state = upf.read_restart(path)
    assert len(state["type"]) == len(typ)
    b = upf.Simulation(p, len(typ))
    b.upload_restart(state)
    b.step(4, method)
    kinds = (UPF_TEXT_POSITIONS, UPF_TEXT_RESTART)
    fa = sim_files(a, tmp_path, "a", kinds)
    fb = sim_files(b, tmp_path, "b", kinds)
    for which in kinds:
        assert fa[which] == fb[which], which
    if method == upf.UPF_PCISPH_STATIC:
        ka, ra = a.stats()
        kb, rb = b.stats()
        assert np.array_equal(ka[-4:], kb) and np.array_equal(ra[-4:], rb)  # stats count from the upload
        # the oracle restarted from the same file gives the same files
        o = ol.OracleSim(p, state["pos"], state["vel"], state["type"])
        o.set_restart(state["corr_pressure"], state["predicted_density"])
        ids = oracle_steps(o, 4, method)
        fo = oracle_files(o, ids, tmp_path, "o", kinds)
        for which in kinds:
            assert fb[which] == fo[which], which


def test_state_file_call_errors(gpu, tmp_path):
    p, pos, typ, nf = scenes.dam_break(8)
    sim = upf.Simulation(p, len(typ))
    L = upf.lib()
    assert L.upf_save_text(sim.h, str(tmp_path / "x.txt").encode(), UPF_TEXT_LIQUID) == UPF_ERR_STATE
    sim.upload(pos, typ)
    assert L.upf_save_text(sim.h, str(tmp_path / "x.txt").encode(), 0) == UPF_ERR_ARG
    bad = typ.copy()
    bad[0] = 5
    z = np.zeros(len(typ), np.float32)
    pos4 = np.ascontiguousarray(pos, np.float32)
    assert L.upf_upload_restart(sim.h, len(typ), upf.ptr(pos4), None, upf.ptr(bad, C.c_int), upf.ptr(z),
                                upf.ptr(z)) == UPF_ERR_ARG
    assert L.upf_upload_restart(sim.h, len(typ), upf.ptr(pos4), None, upf.ptr(typ, C.c_int), None,
                                None) == UPF_ERR_ARG
    # before the first step the files hold the upload order
    sim.save_text(str(tmp_path / "p.txt"), UPF_TEXT_POSITIONS)
    lines = read_bytes(tmp_path / "p.txt").split(b"\n")
    assert int(lines[0]) == len(typ)
    first = np.float32([float(v) for v in lines[1].split()[:3]])
    assert np.array_equal(first, pos4[0, :3])


@pytest.mark.parametrize("steps", [0, 3])
def test_cell_dump_equals_oracle_and_leaves_the_run_unchanged(gpu, tmp_path, steps):
    """upf_dump_cells (DumpGPUParamsToFile) after `steps` steps: the same
    bytes as the oracle's dump of its state after the same steps; the run
    that dumped continues exactly as one that did not."""
    p, pos, typ, nf = scenes.dam_break(16)
    a = upf.Simulation(p, len(typ))
    a.upload(pos, typ)
    if steps:
        a.step(steps, upf.UPF_PCISPH_STATIC)
    gp, gb = str(tmp_path / "g_particles.bin"), str(tmp_path / "g_blocks.bin")
    a.dump_cells(gp, gb)
    o = ol.OracleSim(p, pos, None, typ)
    if steps:
        o.step(steps, upf.UPF_PCISPH_STATIC)
    op, ob = str(tmp_path / "o_particles.bin"), str(tmp_path / "o_blocks.bin")
    o.dump_cells(op, ob)
    assert read_bytes(gp) == read_bytes(op)
    assert read_bytes(gb) == read_bytes(ob)
    a.dump_cells(gp, gb)  # a second dump of the same state appends the same record
    recs = upf.read_cell_dump(gp, gb)
    assert len(recs) == 2 and recs[0][0].tobytes() == recs[1][0].tobytes()
    assert recs[0][1].tobytes() == recs[1][1].tobytes()
    a.step(3, upf.UPF_PCISPH_STATIC)
    b = upf.Simulation(p, len(typ))
    b.upload(pos, typ)
    b.step(steps + 3, upf.UPF_PCISPH_STATIC)
    fa = sim_files(a, tmp_path, "a", (UPF_TEXT_RESTART,))
    fb = sim_files(b, tmp_path, "b", (UPF_TEXT_RESTART,))
    assert fa[UPF_TEXT_RESTART] == fb[UPF_TEXT_RESTART]
