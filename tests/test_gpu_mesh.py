"""OBJ model -> particles on the GPU (upf_mesh_fill) against the oracle and
against the reference's results kept in tests/golden/obj_models.json: the
same particles, bit for bit, in the same order."""
import hashlib
import json
import os

import numpy as np
import pytest

import unifiedparticleframeworkcuda_amd as upf
import oracle_lib as ol

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
with open(os.path.join(GOLD, "obj_models.json")) as _f:
    GOLDEN = json.load(_f)


def _digest(pos):
    return hashlib.sha256(np.ascontiguousarray(pos, np.float32).tobytes()).hexdigest()


@pytest.mark.parametrize("case", GOLDEN["meshes"], ids=lambda c: "%s-%s-%g" % (c["file"], c["scale"][0], c["spacing"]))
def test_gpu_fill_equals_reference_results(gpu, case):
    faces = upf.load_obj(os.path.join(GOLD, case["file"]))
    print("mesh fill %s: %d faces, scale %s, spacing %g" % (case["file"], len(faces), case["scale"], case["spacing"]),
          flush=True)
    pos = upf.mesh_fill(faces, case["scale"], case["spacing"])
    assert len(pos) == case["count"]
    assert _digest(pos) == case["sha256"]


@pytest.mark.parametrize("name,scale,spacing", [("mesh_torus.obj", (1.0, 3.0, 1.5), 0.009),
                                                ("mesh_blob.obj", (2.0, 2.0, 2.0), 0.013),
                                                ("mesh_sphere.obj", (0.5, 0.5, 0.5), 0.002)])
def test_gpu_fill_equals_oracle(gpu, name, scale, spacing):
    faces = upf.load_obj(os.path.join(GOLD, name))
    print("mesh fill vs oracle %s, scale %s, spacing %g" % (name, scale, spacing), flush=True)
    got = upf.mesh_fill(faces, scale, spacing)
    ref = ol.mesh_fill(faces, scale, spacing)
    assert len(got) > 0 and np.array_equal(got, ref)


def test_gpu_fill_degenerate_faces_and_edge_cases(gpu):
    """Zero-area faces (a zero normal), a single triangle (flat: only its
    touching lattice points), an empty face list, count-only and a short
    output buffer."""
    import ctypes as C
    rng = np.random.RandomState(11)
    faces = rng.uniform(-0.05, 0.05, size=(40, 3, 3)).astype(np.float32)
    faces[::7, 1] = faces[::7, 0]  # every seventh face degenerate
    for sc, sp in (((1.0, 1.0, 1.0), 0.01), ((2.0, 1.0, 0.5), 0.004)):
        assert np.array_equal(upf.mesh_fill(faces, sc, sp), ol.mesh_fill(faces, sc, sp))
    tri = np.array([[[0.0, 0.0, 0.0], [0.1, 0.0, 0.0], [0.0, 0.1, 0.0]]], np.float32)
    assert np.array_equal(upf.mesh_fill(tri, 1.0, 0.01), ol.mesh_fill(tri, 1.0, 0.01))
    assert len(upf.mesh_fill(np.zeros((0, 3, 3), np.float32), 1.0, 0.01)) == 0
    L = upf.lib()
    f = np.ascontiguousarray(faces.reshape(-1, 9))
    sc = (C.c_float * 3)(1.0, 1.0, 1.0)
    n = C.c_int()
    assert L.upf_mesh_fill(upf.ptr(f), len(f), sc, 0.01, None, 0, C.byref(n)) == upf.UPF_OK
    out = np.zeros((max(n.value - 1, 1), 4), np.float32)
    assert L.upf_mesh_fill(upf.ptr(f), len(f), sc, 0.01, upf.ptr(out), n.value - 1, C.byref(n)) == \
        upf._abi.UPF_ERR_RANGE
