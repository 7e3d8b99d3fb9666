"""The C ABI as a C++ caller sees it: include/upf.h compiled with g++.

* Layout: a program generated from the ctypes mirror (_abi.py) prints
  offsetof / sizeof of every field of every struct of the header, as the
  compiler lays them out; they must equal the mirror's offsets and sizes (a
  field-order slip, a missing or extra field fails here).
* The reference-side binding (tests/abi/binding_example.cpp: the calls of
  UnifiedPhysics::InitCudaVariables / CudaParticlePhysics against upf.h,
  linked to libupf.so): its host part runs without a device and gives the
  mirror's parameters and scene; on the GPU its step equals the mirror's
  step bit for bit."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import unifiedparticleframeworkcuda_amd as upf
from unifiedparticleframeworkcuda_amd._abi import Params, DamBreakDesc, RigidSceneDesc, Body, Tuning

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
PKG = os.path.join(ROOT, "unifiedparticleframeworkcuda_amd")
STRUCTS = {"upf_params": Params, "upf_dam_break_desc": DamBreakDesc, "upf_rigid_scene_desc": RigidSceneDesc,
           "upf_body": Body, "upf_tuning": Tuning}


def _gxx(src, out, link=False):
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-I" + INC, src, "-o", out]
    if link:
        cmd += ["-L" + PKG, "-lupf", "-Wl,-rpath," + PKG]
    subprocess.check_call(cmd)


def test_struct_layout_compiled_from_the_header(tmp_path):
    lines = ['#include <cstddef>', '#include <cstdio>', '#include "upf.h"', "int main() {"]
    for cname, cls in STRUCTS.items():
        lines.append('std::printf("%s sizeof %%zu\\n", sizeof(%s));' % (cname, cname))
        for field, _ in cls._fields_:
            lines.append('std::printf("%s %s %%zu %%zu\\n", offsetof(%s, %s), sizeof(((%s*)0)->%s));'
                         % (cname, field, cname, field, cname, field))
    lines.append("return 0; }")
    src = tmp_path / "layout.cpp"
    src.write_text("\n".join(lines) + "\n")
    exe = str(tmp_path / "layout")
    _gxx(str(src), exe)
    out = subprocess.check_output([exe], text=True).split("\n")
    got = {}
    for ln in out:
        parts = ln.split()
        if len(parts) == 3 and parts[1] == "sizeof":
            got[(parts[0], None)] = (int(parts[2]),)
        elif len(parts) == 4:
            got[(parts[0], parts[1])] = (int(parts[2]), int(parts[3]))
    for cname, cls in STRUCTS.items():
        assert got[(cname, None)] == (C.sizeof(cls),), cname
        for field, _ in cls._fields_:
            f = getattr(cls, field)
            assert got[(cname, field)] == (f.offset, f.size), (cname, field)


@pytest.fixture(scope="module")
def binding(tmp_path_factory):
    upf.lib()  # built
    exe = str(tmp_path_factory.mktemp("abi") / "binding_example")
    _gxx(os.path.join(ROOT, "tests", "abi", "binding_example.cpp"), exe, link=True)
    return exe


def _mirror_scene():
    return upf.standard_dam_break(12, 12, 12)


def test_binding_host_part(binding):
    out = subprocess.check_output([binding, "host"], text=True).split("\n")
    p, pos, typ, nf = _mirror_scene()
    words = out[0].split()
    assert words[0] == "params"
    vals = dict(zip(words[1::2], words[2::2]))
    assert np.float32(vals["spacing"]) == np.float32(p.spacing)
    assert np.float32(vals["h"]) == np.float32(p.support_radius)
    assert np.float32(vals["mass"]) == np.float32(p.initial_mass)
    assert np.float32(vals["delta"]) == np.float32(p.density_error_factor)
    assert int(vals["block"]) == p.block_size and int(vals["grid"]) == p.grid_dim
    assert out[1].split() == ["scene", "fluid", str(nf), "wall", str(len(typ) - nf)]


@pytest.mark.gpu
def test_binding_step_equals_the_mirror(gpu, binding):
    out = subprocess.check_output([binding, "step", "2"], text=True).strip().split("\n")[2:]
    got = np.array([[int(w, 16) for w in ln.split()] for ln in out], np.uint32).view(np.float32)
    p, pos, typ, nf = _mirror_scene()
    sim = upf.Simulation(p, len(typ))
    sim.upload(pos, typ)
    sim.step(2, upf.UPF_PCISPH_STATIC)
    exp = sim.state()["pos"][:, :3]
    assert got.shape == exp.shape
    assert np.array_equal(got, exp)
