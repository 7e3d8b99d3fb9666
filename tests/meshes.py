"""Procedural triangle meshes written as Wavefront OBJ text in the format the
reference's ObjLoader reads (faces as v/vt/vn triples), for the model-fill
tests: a UV sphere, a torus, a bumpy (non-convex) blob, an axis-aligned box
and a mesh with degenerate faces."""
import numpy as np


def _obj(verts, tris, trailer="comment"):
    lines = ["# procedural mesh", "# %d vertices, %d faces" % (len(verts), len(tris))]
    lines += ["v %.6f %.6f %.6f" % tuple(v) for v in verts]
    lines += ["vt 0.0 0.0", "vn 0.0 1.0 0.0"]
    lines += ["f %d/1/1 %d/1/1 %d/1/1" % (a + 1, b + 1, c + 1) for a, b, c in tris]
    if trailer == "comment":
        lines.append("# end")
    text = "\n".join(lines)
    # trailer "face": the file ends with a face line and its newline, which
    # the reference's reader processes twice
    return text + "\n"


def _grid_tris(nu, nv, wrap_v=True):
    tris = []
    for i in range(nu):
        for j in range(nv if wrap_v else nv - 1):
            a = i * nv + j
            b = ((i + 1) % nu) * nv + j
            c = ((i + 1) % nu) * nv + (j + 1) % nv
            d = i * nv + (j + 1) % nv
            tris += [(a, b, c), (a, c, d)]
    return tris


def sphere(r=0.1, nu=24, nv=12, center=(0.0, 0.0, 0.0), bump=0.0, seed=0, trailer="comment"):
    """UV sphere (poles as rings of coincident vertices: their faces are
    degenerate); bump > 0 perturbs the radius by a smooth seeded field."""
    rng = np.random.RandomState(seed)
    k = rng.uniform(1.0, 3.0, size=3)
    ph = rng.uniform(0.0, 2 * np.pi, size=3)
    verts = []
    for i in range(nu):
        t = 2 * np.pi * i / nu
        for j in range(nv):
            p = np.pi * j / (nv - 1)
            d = np.array([np.sin(p) * np.cos(t), np.cos(p), np.sin(p) * np.sin(t)])
            rr = r * (1.0 + bump * np.sin(k[0] * t + ph[0]) * np.sin(k[1] * p + ph[1]) * np.cos(k[2] * (t + p) + ph[2]))
            verts.append(np.asarray(center) + rr * d)
    return _obj(np.array(verts), _grid_tris(nu, nv, wrap_v=False), trailer)


def torus(R=0.12, r=0.04, nu=32, nv=16, center=(0.0, 0.0, 0.0), trailer="comment"):
    verts = []
    for i in range(nu):
        t = 2 * np.pi * i / nu
        for j in range(nv):
            p = 2 * np.pi * j / nv
            verts.append(np.asarray(center) + [(R + r * np.cos(p)) * np.cos(t), r * np.sin(p),
                                               (R + r * np.cos(p)) * np.sin(t)])
    return _obj(np.array(verts), _grid_tris(nu, nv), trailer)


def box(a=(0.2, 0.1, 0.15), lo=(0.0, 0.0, 0.0), trailer="comment"):
    """Axis-aligned box [lo, lo + a], 12 triangles."""
    lo = np.asarray(lo, float)
    a = np.asarray(a, float)
    c = [lo + a * np.array([(k >> 0) & 1, (k >> 1) & 1, (k >> 2) & 1]) for k in range(8)]
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    tris = []
    for q in quads:
        tris += [(q[0], q[1], q[2]), (q[0], q[2], q[3])]
    return _obj(np.array(c), tris, trailer)


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)
