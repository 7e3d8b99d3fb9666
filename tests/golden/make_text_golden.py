"""Writes tests/golden/state_files.json: particle arrays, the bytes the
reference's UnifiedIO writers (src/cpu/UnifiedIO.cpp:13-166) produce for
them, and a restart file with the values the reference's reader
(ReadInFluidParticlesPCISPH, :196-220) reads from it.  Needs oracle/_ref built
from the reference tree (make -C oracle ref); run from the repository root:
    python tests/golden/make_text_golden.py
"""
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle_lib as ol  # noqa: E402


def particles(n=48, seed=11):
    rng = np.random.default_rng(seed)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.normal(size=(n, 3)) * 10.0 ** rng.integers(-6, 6, size=(n, 3))
    # values whose printing exercises rounding, padding and exponent forms
    edge = [0.0, -0.0, 1e-45, -1.17549435e-38, 3.4028235e38, 0.1, 1.0 / 3.0, 123456789.0, 1e-5, 99999999.0,
            0.99999999, 1e16, -2.5e-7, 12345678.0, 1.5, -1.0]
    flat = pos[:, :3].reshape(-1)
    flat[:len(edge)] = np.asarray(edge, np.float32)
    pos[:, :3] = flat.reshape(-1, 3)
    types = rng.choice([0, 2, 6], size=n).astype(np.int32)
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = rng.normal(size=(n, 3))
    cp = (rng.random(n) * 3000.0).astype(np.float32)
    rho = (1000.0 + rng.normal(size=n) * 5.0).astype(np.float32)
    return pos, types, vel, cp, rho


# A restart file in the forms an istream reads: several particles on one line,
# tabs, CRLF, signs, exponents in both cases, leading zeros, ".5" and "5."
RESTART_TEXT = ("3\r\n"
                "0.5 -1.25 +2e-3 2 1E+1 -0 .5 12.75 1000.125\r\n"
                "\t5. 007 -3.4028235e+38\t6  0 0 0 0 999.999\n"
                "1e-45 -1.17549435E-38 0.1 2 -1.5 2.5 -3.5 0.000123 1001 trailing text is not read\n")


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32).reshape(-1).tolist()


def main():
    if ol.ref_lib() is None:
        sys.exit("oracle/_ref is not built (needs the reference tree)")
    pos, types, vel, cp, rho = particles()
    out = {"n": len(types), "pos_bits": bits(pos[:, :3]), "type": types.tolist(), "vel_bits": bits(vel[:, :3]),
           "corr_pressure_bits": bits(cp), "predicted_density_bits": bits(rho)}
    with tempfile.TemporaryDirectory() as d:
        for which, name in ((1, "liquid"), (2, "rigid"), (3, "positions")):
            out[name] = ol.ref_save_text(d, which, pos, types).decode("ascii")
        path = os.path.join(d, "restart.txt")
        with open(path, "w", newline="") as f:
            f.write(RESTART_TEXT)
        n, r = ol.ref_read_restart(path)
    out["restart_text"] = RESTART_TEXT
    out["restart_expected"] = {"n": n, "pos_bits": bits(r["pos"]), "vel_bits": bits(r["vel"]),
                               "type": r["type"].tolist(), "corr_pressure_bits": bits(r["corr_pressure"]),
                               "predicted_density_bits": bits(r["predicted_density"])}
    with open(os.path.join(HERE, "state_files.json"), "w") as f:
        json.dump(out, f, indent=0)
    print("wrote state_files.json:", {k: len(v) for k, v in out.items() if isinstance(v, str)})


if __name__ == "__main__":
    main()
