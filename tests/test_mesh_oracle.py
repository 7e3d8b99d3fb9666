"""OBJ model -> particles (ObjLoader / ObjectLoader::LoadParticles, row "scene
input"): the oracle against the reference's own translation units (compiled
into oracle/_ref, and their results kept in tests/golden/obj_models.json),
against the geometry of an axis-aligned box, and the library's OBJ reader
(host code) against the oracle's.  No GPU needed."""
import hashlib
import json
import os

import numpy as np
import pytest

import unifiedparticleframeworkcuda_amd as upf
import oracle_lib as ol

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
MODELS = "/root/reference/models"
with open(os.path.join(GOLD, "obj_models.json")) as _f:
    GOLDEN = json.load(_f)


def _digest(pos):
    return hashlib.sha256(np.ascontiguousarray(pos, np.float32).tobytes()).hexdigest()


@pytest.mark.parametrize("case", GOLDEN["meshes"], ids=lambda c: "%s-%s-%g" % (c["file"], c["scale"][0], c["spacing"]))
def test_oracle_fill_equals_reference_on_fixture_meshes(case):
    faces = ol.obj_read(os.path.join(GOLD, case["file"]))
    pos = ol.mesh_fill(faces, case["scale"], case["spacing"])
    assert len(pos) == case["count"]
    assert _digest(pos) == case["sha256"]
    assert np.array_equal(pos[:3], np.asarray(case["first"], np.float32))


@pytest.mark.parametrize("case", GOLDEN["reference_models"], ids=lambda c: c["file"])
def test_oracle_fill_equals_reference_on_its_models(case):
    path = os.path.join(MODELS, case["file"])
    if not os.path.exists(path):
        pytest.skip("the reference's models are not in this environment")
    pos = ol.mesh_fill(ol.obj_read(path), case["scale"], case["spacing"])
    assert len(pos) == case["count"] and _digest(pos) == case["sha256"]


def test_oracle_fill_equals_reference_build_directly():
    """Fresh configurations against the reference's compiled loader (when the
    reference tree is here)."""
    if ol.ref_lib() is None:
        pytest.skip("oracle/_ref not built (no reference tree)")
    for name, sc, sp in (("mesh_torus.obj", (1.0, 3.0, 1.5), 0.009), ("mesh_blob.obj", (2.0, 2.0, 2.0), 0.013)):
        path = os.path.join(GOLD, name)
        ref = ol.ref_obj_particles(path, sc, sp)
        got = ol.mesh_fill(ol.obj_read(path), sc, sp)
        assert np.array_equal(ref, got), name


def test_box_fill_matches_the_box_where_the_crossing_is_clean():
    """Axis-aligned box [0, a]: no kept point lies outside the box (beyond
    1e-4), and every lattice point strictly inside whose segment from the
    common start point o = (-spacing,)*3 enters the box through the interior
    of a triangle -- at least 1e-3 away from every face edge and from the
    diagonal that splits each face -- is kept.  (The reference's crossing
    parity is not robust where the segment meets an edge or a corner; those
    points are its numerics' to decide.)"""
    a = np.array([0.2, 0.1, 0.15])
    sp = 0.01
    pos = ol.mesh_fill(ol.obj_read(os.path.join(GOLD, "mesh_box.obj")), (1.0, 1.0, 1.0), sp).astype(np.float64)
    assert np.all(pos >= -1e-4) and np.all(pos <= a + 1e-4)
    kept = {tuple(np.round(p, 5)) for p in pos}
    o = np.full(3, -sp)
    checked = 0
    for x in range(0, 21):
        off = 0.5 * sp if x % 2 else 0.0
        for y in range(0, 12):
            for z in range(0, 17):
                p = np.array([x * sp, off + y * sp, off + z * sp])
                if not np.all((p > 1e-4) & (p < a - 1e-4)):
                    continue
                d = p - o
                t = (0.0 - o) / d  # entry through the faces at 0 (o lies below every face at 0)
                k = int(np.argmax(t))
                q = o + t[k] * d
                u, v = [i for i in range(3) if i != k]
                qu, qv = q[u] / a[u], q[v] / a[v]  # position on the face, in [0, 1]^2
                clean = min(qu, qv, 1 - qu, 1 - qv) * min(a[u], a[v]) > 1e-3
                clean &= abs(qu - qv) * min(a[u], a[v]) > 1e-3 and abs(qu + qv - 1) * min(a[u], a[v]) > 1e-3
                if clean:
                    checked += 1
                    assert tuple(np.round(p, 5)) in kept, p
    assert checked > 500


def test_reader_duplicates_the_last_line_like_the_reference():
    """mesh_box.obj ends with a face line: 12 faces written, 13 read."""
    f = ol.obj_read(os.path.join(GOLD, "mesh_box.obj"))
    assert len(f) == 13 and np.array_equal(f[12], f[11])
    g = ol.obj_read(os.path.join(GOLD, "mesh_sphere.obj"))  # ends with a comment
    assert len(g) == 2 * 24 * 11


@pytest.mark.parametrize("name", ["mesh_sphere.obj", "mesh_blob.obj", "mesh_torus.obj", "mesh_box.obj"])
def test_library_reader_equals_oracle_reader(name):
    path = os.path.join(GOLD, name)
    assert np.array_equal(upf.load_obj(path), ol.obj_read(path))


def test_library_reader_errors(tmp_path):
    with pytest.raises(upf.UPFError) as e:
        upf.load_obj(str(tmp_path / "missing.obj"))
    assert e.value.code == upf._abi.UPF_ERR_IO
    p = tmp_path / "plain_faces.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")  # no v/vt/vn triples
    with pytest.raises(upf.UPFError) as e:
        upf.load_obj(str(p))
    assert e.value.code == upf._abi.UPF_ERR_ARG
    q = tmp_path / "no_faces.obj"
    q.write_text("v 0 0 0\nv 1 0 0\n")
    assert len(upf.load_obj(str(q))) == 0  # the reference's FileError: an empty model
