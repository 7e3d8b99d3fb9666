"""Writes tests/golden/reference_units.json from oracle/_ref/libupfref.so,
i.e. from the reference's own translation units (WeightingKernel.cpp,
UnifiedConstants.cpp, RigidBody.cpp, Matrix3x3.cpp, UnifiedParticle.cpp)
compiled in this container by oracle/Makefile.  The file holds inputs and
expected outputs only (float bit patterns).

    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle_lib as ol  # noqa: E402
from unifiedparticleframeworkcuda_amd._abi import ptr  # noqa: E402


def main():
    ref = ol.ref_lib()
    if ref is None:
        raise SystemExit("oracle/_ref/libupfref.so missing: build it with `make -C oracle ref`")
    out = {"source": "oracle/_ref (reference translation units)", "luts": [], "constants": [], "rigid": []}
    for h in (0.0200002, 0.0500005, 0.0100001):
        h = float(np.float32(h))
        L = np.zeros(1505, np.float32)
        ks = np.zeros(1, np.float32)
        z = np.zeros(1, np.float32)
        ref.ref_luts(h, ptr(L), ptr(ks), 0, ptr(z), ptr(z), ptr(z), ptr(z))
        out["luts"].append({"h": h, "bits": L.view(np.uint32).tolist(),
                            "kernel_self_bits": int(ks.view(np.uint32)[0])})
    for sp in (0.01, 0.025, 0.02):
        c = np.zeros(32, np.float32)
        ref.ref_constants(sp, ptr(c))
        out["constants"].append({"spacing": sp, "bits": c[:20].view(np.uint32).tolist()})
    rng = np.random.RandomState(7)
    for trial in range(4):
        st = np.zeros(47, np.float32)
        st[0:3] = rng.uniform(0.2, 0.8, 3)
        st[9:12] = rng.normal(0, 0.05, 3)
        st[12:15] = rng.normal(0, 1e-4, 3)
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        st[15:19] = q
        w, x, y, zq = q
        R = np.array([[1 - 2 * (y * y + zq * zq), 2 * (x * y - w * zq), 2 * (x * zq + w * y)],
                      [2 * (x * y + w * zq), 1 - 2 * (x * x + zq * zq), 2 * (y * zq - w * x)],
                      [2 * (x * zq - w * y), 2 * (y * zq + w * x), 1 - 2 * (x * x + y * y)]])
        st[19:28] = R.ravel()
        st[28:37] = np.linalg.inv(np.diag(rng.uniform(1e-4, 1e-3, 3))).ravel()
        st[46] = rng.uniform(0.5, 3.0)
        n = 20
        pos = (st[0:3] + rng.uniform(-0.05, 0.05, (n, 3))).astype(np.float32)
        f = rng.normal(0, 0.01, (n, 3)).astype(np.float32)
        m = np.full(n, st[46] / n, np.float32)
        a = st.copy()
        for _ in range(3):
            ref.ref_rigid_step(ptr(a), n, ptr(pos), ptr(f), ptr(m), 0.001, 9.81, 20.0)
        out["rigid"].append({"steps": 3, "state_in": st.view(np.uint32).tolist(), "pos": pos.view(np.uint32).ravel().tolist(),
                             "force": f.view(np.uint32).ravel().tolist(), "mass": m.view(np.uint32).tolist(),
                             "state_out": a.view(np.uint32).tolist()})
    with open(os.path.join(HERE, "reference_units.json"), "w") as fh:
        json.dump(out, fh)
    print("wrote", os.path.join(HERE, "reference_units.json"))


if __name__ == "__main__":
    main()
