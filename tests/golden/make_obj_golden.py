"""Writes the OBJ model-fill fixtures: four procedural meshes
(tests/golden/mesh_*.obj, made here by tests/meshes.py) and
tests/golden/obj_models.json -- particle counts and SHA-256 of the float32
positions that the reference's own ObjLoader / ObjectLoader / Polygon
translation units (compiled into oracle/_ref by oracle/Makefile) produce for
those meshes and for the three models shipped with the reference.

    python tests/golden/make_obj_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import meshes  # noqa: E402
import oracle_lib as ol  # noqa: E402

MESHES = {
    "mesh_sphere.obj": lambda: meshes.sphere(0.1, 24, 12, center=(0.03, 0.12, -0.02)),
    "mesh_blob.obj": lambda: meshes.sphere(0.08, 40, 21, center=(-0.05, 0.1, 0.04), bump=0.25, seed=3,
                                           trailer="face"),
    "mesh_torus.obj": lambda: meshes.torus(0.12, 0.04, 32, 16, center=(0.0, 0.06, 0.0)),
    "mesh_box.obj": lambda: meshes.box((0.2, 0.1, 0.15), lo=(0.0, 0.0, 0.0), trailer="face"),
}
# (mesh, scale, spacing): the reference's bunny placement (scale 2.5, spacing
# 0.01), a coarse and a fine lattice, anisotropic scales
CONFIGS = [("mesh_sphere.obj", (2.5, 2.5, 2.5), 0.01), ("mesh_sphere.obj", (1.0, 1.0, 1.0), 0.007),
           ("mesh_blob.obj", (2.5, 2.5, 2.5), 0.01), ("mesh_blob.obj", (1.5, 2.0, 1.0), 0.005),
           ("mesh_torus.obj", (2.0, 2.0, 2.0), 0.01), ("mesh_box.obj", (1.0, 1.0, 1.0), 0.01),
           ("mesh_box.obj", (3.0, 3.0, 3.0), 0.02)]
REFERENCE_MODELS = [("bunny-1500.obj", (2.5, 2.5, 2.5), 0.01), ("teapot.obj", (0.00025, 0.00025, 0.00025), 0.01),
                    ("banana.obj", (0.02, 0.02, 0.02), 0.01)]
MODELS_DIR = "/root/reference/models"


def digest(pos):
    return hashlib.sha256(np.ascontiguousarray(pos, np.float32).tobytes()).hexdigest()


def main():
    if ol.ref_lib() is None:
        raise SystemExit("oracle/_ref/libupfref.so missing: build it with `make -C oracle ref`")
    for name, make in MESHES.items():
        with open(os.path.join(HERE, name), "w") as f:
            f.write(make())
    out = {"source": "oracle/_ref (reference ObjLoader.cpp, ObjectLoader.cpp, Polygon.cpp, Face.cpp, Mesh.cpp)",
           "meshes": [], "reference_models": []}
    for name, sc, sp in CONFIGS:
        pos = ol.ref_obj_particles(os.path.join(HERE, name), sc, sp)
        out["meshes"].append({"file": name, "scale": list(sc), "spacing": sp, "count": int(len(pos)),
                              "sha256": digest(pos), "first": pos[:3].tolist()})
        print(name, sc, sp, len(pos))
    for name, sc, sp in REFERENCE_MODELS:
        pos = ol.ref_obj_particles(os.path.join(MODELS_DIR, name), sc, sp)
        out["reference_models"].append({"file": name, "scale": list(sc), "spacing": sp, "count": int(len(pos)),
                                        "sha256": digest(pos)})
        print(name, sc, sp, len(pos))
    with open(os.path.join(HERE, "obj_models.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
