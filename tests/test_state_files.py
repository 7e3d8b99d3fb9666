"""Particle state files (src/cpu/UnifiedIO.cpp): the writers' bytes and the
restart reader's values.  The oracle (iostream formatting, as the reference)
is pinned against golden bytes written by the reference's own UnifiedIO.cpp
(tests/golden/make_text_golden.py) and, where oracle/_ref is built, against
that library directly; the library's writers (printf formatting) and reader
(strtof) are host code, checked here against both without a device."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle_lib as ol
import unifiedparticleframeworkcuda_amd as upf
from unifiedparticleframeworkcuda_amd import UPF_TEXT_LIQUID, UPF_TEXT_RIGID, UPF_TEXT_POSITIONS, UPF_TEXT_RESTART
from unifiedparticleframeworkcuda_amd._abi import UPF_ERR_ARG, UPF_ERR_IO, UPF_ERR_RANGE, ptr

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "state_files.json")
KINDS = ((UPF_TEXT_LIQUID, "liquid"), (UPF_TEXT_RIGID, "rigid"), (UPF_TEXT_POSITIONS, "positions"))


def f32(bits, shape=None):
    a = np.asarray(bits, np.uint32).view(np.float32)
    return a.reshape(shape) if shape else a


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        g = json.load(f)
    n = g["n"]
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = f32(g["pos_bits"], (n, 3))
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = f32(g["vel_bits"], (n, 3))
    g["arrays"] = dict(pos=pos, type=np.asarray(g["type"], np.int32), vel=vel,
                       cp=f32(g["corr_pressure_bits"]), rho=f32(g["predicted_density_bits"]))
    return g


def read_bytes(path):
    with open(path, "rb") as f:
        return f.read()


@pytest.mark.parametrize("which,name", KINDS)
def test_oracle_writer_equals_reference_bytes(golden, tmp_path, which, name):
    a = golden["arrays"]
    ol.write_text(str(tmp_path / "o.txt"), which, a["pos"], a["type"])
    assert read_bytes(tmp_path / "o.txt") == golden[name].encode()


@pytest.mark.parametrize("which,name", KINDS)
def test_library_writer_equals_reference_bytes(golden, tmp_path, which, name):
    a = golden["arrays"]
    upf.write_text(str(tmp_path / "u.txt"), which, a["pos"], a["type"])
    assert read_bytes(tmp_path / "u.txt") == golden[name].encode()


def test_golden_files_have_the_reference_shape(golden):
    """Independent of any writer: the count line counts the kept particles,
    every particle line of the liquid / rigid files has the reference's
    field widths and zeros, and positions parse back to the stored floats
    (20 significant digits print a float exactly)."""
    a = golden["arrays"]
    t = a["type"]
    for name, keep in (("liquid", t == 2), ("rigid", (t == 0) | (t == 6)), ("positions", np.ones_like(t, bool))):
        lines = golden[name].split("\n")
        assert lines[-1] == ""
        assert int(lines[0]) == int(keep.sum()) == len(lines) - 2
        for line, i in zip(lines[1:-1], np.nonzero(keep)[0]):
            if name == "positions":
                f = line.split(" ")
                assert f[3:] == ["0"] * 8 + [""]
                assert np.array_equal(np.float32([float(v) for v in f[:3]]).view(np.uint32),
                                      a["pos"][i, :3].view(np.uint32))
            else:
                assert line.endswith("0               0 0 0  0  0  0  ")
                for k, w in enumerate((15, 15, 15)):
                    field = line[16 * k:16 * k + 15]
                    assert len(field) == w and field.rstrip() == field.strip()
                    assert np.float32(float(field)) == np.float32("%.8g" % a["pos"][i, k])
                assert int(line[48:51]) == t[i]


def test_restart_file_round_trips_every_float(golden, tmp_path):
    a = golden["arrays"]
    rng = np.random.default_rng(4)
    n = len(a["type"])
    specials = np.float32([0.0, -0.0, 1e-45, -1e-45, 1.17549435e-38, 3.4028235e38, -3.4028235e38, 1.0000001])
    vel = a["vel"].copy()
    vel[:2, :3] = specials[:6].reshape(2, 3)
    vel[2, :2] = specials[6:]
    bits = rng.integers(0, 0x7F800000, size=n, dtype=np.uint32).view(np.float32)  # every finite positive float
    for write, name in ((upf.write_text, "lib"), (ol.write_text, "orc")):
        path = str(tmp_path / ("r_%s.txt" % name))
        write(path, UPF_TEXT_RESTART, a["pos"], a["type"], vel, bits, a["rho"])
        st = upf.read_restart(path)
        assert np.array_equal(st["pos"][:, :3].view(np.uint32), a["pos"][:, :3].view(np.uint32))
        assert np.array_equal(st["vel"][:, :3].view(np.uint32), vel[:, :3].view(np.uint32))
        assert np.array_equal(st["type"], a["type"])
        assert np.array_equal(st["corr_pressure"].view(np.uint32), bits.view(np.uint32))
        assert np.array_equal(st["predicted_density"].view(np.uint32), a["rho"].view(np.uint32))
        assert np.all(st["pos"][:, 3] == 0) and np.all(st["vel"][:, 3] == 0)
    assert read_bytes(tmp_path / "r_lib.txt") == read_bytes(tmp_path / "r_orc.txt")


def test_readers_equal_reference_reader(golden, tmp_path):
    path = tmp_path / "restart.txt"
    path.write_bytes(golden["restart_text"].encode())
    e = golden["restart_expected"]
    n = e["n"]
    exp = dict(pos=f32(e["pos_bits"], (n, 3)), vel=f32(e["vel_bits"], (n, 3)), type=np.asarray(e["type"], np.int32),
               corr_pressure=f32(e["corr_pressure_bits"]), predicted_density=f32(e["predicted_density_bits"]))
    got = upf.read_restart(str(path))
    cnt, orc = ol.read_restart(str(path))
    assert cnt == n and len(orc["type"]) == n
    for k, v in exp.items():
        lib_v = got[k][:, :3] if k in ("pos", "vel") else got[k]
        assert np.array_equal(np.asarray(lib_v).view(np.uint32), np.asarray(v).view(np.uint32)), k
        assert np.array_equal(np.asarray(orc[k]).view(np.uint32), np.asarray(v).view(np.uint32)), k


def test_reference_library_agrees_on_random_particles(tmp_path):
    """Reference writers and reader compiled from the reference tree (skipped
    without it) against the oracle and the library, on 70,000 particles (the
    library formats them on several threads)."""
    if ol.ref_lib() is None:
        pytest.skip("oracle/_ref not built (no reference tree)")
    rng = np.random.default_rng(9)
    n = 70000
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.normal(size=(n, 3)) * 10.0 ** rng.integers(-9, 9, size=(n, 3))
    t = rng.choice([0, 2, 6], size=n).astype(np.int32)
    for which, _ in KINDS:
        ref = ol.ref_save_text(str(tmp_path), which, pos, t)
        upf.write_text(str(tmp_path / "u.txt"), which, pos, t)
        ol.write_text(str(tmp_path / "o.txt"), which, pos, t)
        assert read_bytes(tmp_path / "u.txt") == ref
        assert read_bytes(tmp_path / "o.txt") == ref
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = rng.normal(size=(n, 3))
    cp = rng.random(n).astype(np.float32)
    rho = rng.random(n).astype(np.float32) + 999
    path = str(tmp_path / "r.txt")
    upf.write_text(path, UPF_TEXT_RESTART, pos, t, vel, cp, rho)
    cnt, r = ol.ref_read_restart(path, cap=n)
    assert cnt == n
    assert np.array_equal(r["pos"].view(np.uint32), pos[:, :3].view(np.uint32))
    assert np.array_equal(r["vel"], vel[:, :3]) and np.array_equal(r["type"], t)
    assert np.array_equal(r["corr_pressure"], cp) and np.array_equal(r["predicted_density"], rho)


def _raw_read(path, cap):
    n = C.c_int(-1)
    bufs = [np.zeros((max(cap, 1), 4), np.float32), np.zeros((max(cap, 1), 4), np.float32),
            np.zeros(max(cap, 1), np.int32), np.zeros(max(cap, 1), np.float32), np.zeros(max(cap, 1), np.float32)]
    rc = upf.lib().upf_read_restart(os.fsencode(path), cap, C.byref(n), ptr(bufs[0]), ptr(bufs[1]),
                                    ptr(bufs[2], C.c_int), ptr(bufs[3]), ptr(bufs[4]))
    return rc, n.value


@pytest.mark.parametrize("text,rc", [
    ("2\n1 2 3 2 0 0 0 0 1000\n", UPF_ERR_ARG),                         # fewer particles than stated
    ("1\n1 2 3 2.5 0 0 0 0 1000\n", UPF_ERR_ARG),                       # type is not an int
    ("1\n1 2 3 2 0 0 0 nan 1000\n", UPF_ERR_ARG),                       # not a decimal number
    ("1\n1 2 0x10 2 0 0 0 0 1000\n", UPF_ERR_ARG),                      # hexadecimal
    ("1\n1 2 1e39 2 0 0 0 0 1000\n", UPF_ERR_ARG),                      # overflows a float
    ("x\n", UPF_ERR_ARG),                                               # no count
    ("-1\n", UPF_ERR_ARG),
])
def test_reader_rejects_malformed_files(tmp_path, text, rc):
    path = tmp_path / "bad.txt"
    path.write_text(text)
    assert _raw_read(str(path), 4)[0] == rc


def test_reader_count_capacity_and_missing_file(tmp_path):
    path = tmp_path / "ok.txt"
    path.write_text("3\n" + "1 2 3 2 0 0 0 0 1000\n" * 3)
    n = C.c_int()
    L = upf.lib()
    assert L.upf_read_restart(os.fsencode(str(path)), 0, C.byref(n), None, None, None, None, None) == 0
    assert n.value == 3
    assert _raw_read(str(path), 2) == (UPF_ERR_RANGE, 3)
    assert _raw_read(str(path), 3) == (0, 3)
    assert _raw_read(str(tmp_path / "missing.txt"), 3)[0] == UPF_ERR_IO
    # underflow keeps the rounded value, as an istream does
    path.write_text("1\n1e-50 2 3 2 0 0 0 0 1000\n")
    st = upf.read_restart(str(path))
    assert st["pos"][0, 0] == 0.0 and st["pos"][0, 1] == 2.0


def test_writer_argument_errors(tmp_path):
    L = upf.lib()
    p4 = np.zeros((1, 4), np.float32)
    t = np.full(1, 2, np.int32)
    assert L.upf_write_text(os.fsencode(str(tmp_path / "a.txt")), 9, 1, ptr(p4), ptr(t, C.c_int), None, None,
                            None) == UPF_ERR_ARG
    assert L.upf_write_text(os.fsencode(str(tmp_path / "a.txt")), UPF_TEXT_RESTART, 1, ptr(p4), ptr(t, C.c_int),
                            None, None, None) == UPF_ERR_ARG
    assert L.upf_write_text(os.fsencode(str(tmp_path / "no" / "dir.txt")), UPF_TEXT_LIQUID, 1, ptr(p4),
                            ptr(t, C.c_int), None, None, None) == UPF_ERR_IO
    upf.write_text(str(tmp_path / "empty.txt"), UPF_TEXT_LIQUID, np.zeros((0, 4), np.float32), np.zeros(0, np.int32))
    assert read_bytes(tmp_path / "empty.txt") == b"0\n"


@pytest.mark.parametrize("mode", [upf.UPF_BOUNDARY_CPU_REF, upf.UPF_BOUNDARY_AKINCI])
def test_oracle_restart_continues_exactly(tmp_path, mode):
    """The oracle restarted from its own restart file (positions, velocities,
    correction pressures and predicted densities of the last step) continues
    as the run that wrote it: the pressure of two steps back, which the file
    does not hold, does not reach the state."""
    import scenes
    p, pos, typ, nf = scenes.dam_break(8, mode=mode)
    a = ol.OracleSim(p, pos, None, typ, omp=False)
    a.step(3, upf.UPF_PCISPH_STATIC)
    st = a.slot_state()
    path = str(tmp_path / "r.txt")
    ol.write_text(path, UPF_TEXT_RESTART, st["pos"], st["type"], st["vel"], st["p"], st["rho"])
    a.step(4, upf.UPF_PCISPH_STATIC)
    cnt, r = ol.read_restart(path)
    assert cnt == len(typ)
    p4 = np.zeros((cnt, 4), np.float32)
    p4[:, :3] = r["pos"]
    v4 = np.zeros((cnt, 4), np.float32)
    v4[:, :3] = r["vel"]
    b = ol.OracleSim(p, p4, v4, r["type"], omp=False)
    b.set_restart(r["corr_pressure"], r["predicted_density"])
    b.step(4, upf.UPF_PCISPH_STATIC)
    sa, sb = a.slot_state(), b.slot_state()
    for k in ("pos", "vel", "p", "rho", "type"):
        assert np.array_equal(sa[k].view(np.uint32), sb[k].view(np.uint32)), k
    assert np.array_equal(a.stats()[0][-4:], b.stats()[0])


def test_reader_on_large_files_with_any_whitespace(tmp_path):
    """Files above 1 MiB are parsed by several threads (the text cut at
    whitespace): fields spread over lines in any way, runs of mixed
    whitespace, trailing tokens after the last particle, against the oracle's
    istream reader; a malformed field near the end is still an error."""
    rng = np.random.default_rng(21)
    n = 60000
    vals = rng.normal(size=(n, 9)).astype(np.float32) * np.float32(1000)
    vals[:, 3] = rng.choice([0, 2, 6], size=n)
    toks = ["%d" % n]
    for row in vals:
        for k, v in enumerate(row):
            toks.append("%d" % int(v) if k == 3 else repr(float(v)))
    toks += ["not", "read"]
    seps = rng.choice([" ", "\t", "\n", "  \r\n", " \t "], size=len(toks))
    text = "".join(t + s for t, s in zip(toks, seps))
    assert len(text) > (1 << 20)
    path = tmp_path / "big.txt"
    path.write_text(text)
    got = upf.read_restart(str(path))
    cnt, orc = ol.read_restart(str(path), cap=n)
    assert cnt == n and len(orc["type"]) == n
    assert np.array_equal(got["pos"][:, :3], orc["pos"]) and np.array_equal(got["vel"][:, :3], orc["vel"])
    assert np.array_equal(got["type"], orc["type"])
    assert np.array_equal(got["corr_pressure"], orc["corr_pressure"])
    assert np.array_equal(got["predicted_density"], orc["predicted_density"])
    assert np.array_equal(got["pos"][:, :3], vals[:, :3]) and np.array_equal(got["predicted_density"], vals[:, 8])
    toks[-5] = "1.0.0"  # a field of the last particle
    path.write_text("".join(t + s for t, s in zip(toks, seps)))
    assert _raw_read(str(path), n)[0] == UPF_ERR_ARG
    toks[-5] = "1.0"
    toks = toks[:-3]  # the last particle is one field short
    path.write_text(" ".join(toks))
    assert _raw_read(str(path), n)[0] == UPF_ERR_ARG


def _decimal_tokens(rng, count):
    """Decimal numbers in many forms: printed floats and doubles, random digit
    strings of 1-25 digits with any point position and exponent, float
    midpoints (exact and perturbed in the last digit), leading zeros, signs,
    subnormals and underflows."""
    from decimal import Decimal
    out = []
    f = rng.integers(0, 0x7F7FFFFF, size=count // 4, dtype=np.uint32).view(np.float32)
    out += ["%.9g" % x for x in f[: count // 8]] + ["%.17g" % x for x in f[count // 8:]]
    for _ in range(count // 4):
        nd = int(rng.integers(1, 26))
        digits = "".join(rng.choice(list("0123456789"), size=nd))
        dot = int(rng.integers(0, nd + 1))
        s = digits[:dot] + "." + digits[dot:] if rng.random() < 0.7 else digits
        if s == ".":
            s = "0."
        if rng.random() < 0.6:
            s += rng.choice(["e", "E"]) + rng.choice(["", "+", "-"]) + str(int(rng.integers(0, 40)))
        out.append(rng.choice(["", "-", "+"]) + s)
    g = rng.integers(0x00800000, 0x7F000000, size=count // 4, dtype=np.uint32).view(np.float32)
    for x in g:
        mid = (float(x) + float(np.nextafter(x, np.float32(np.inf)))) / 2.0
        exact = format(Decimal(mid), "f") if abs(mid) < 1e30 else str(Decimal(mid))
        out.append(exact)
        out.append("%.16e" % mid)
        out.append("%.8e" % mid)
    # decimals m * 10^p (m <= 2^53) whose nearest double is a float midpoint
    # they are not equal to: rounding through the double would round twice
    traps = 0
    while traps < 4000:
        e, k = int(rng.integers(54, 127)), int(rng.integers(1 << 23, 1 << 24))
        mid = (2 * k + 1) << (e - 24)
        for p in range(1, 23):
            m = (mid + 10 ** p // 2) // 10 ** p
            if 0 < m <= 2 ** 53 and m * 10 ** p != mid and int(float(m * 10 ** p)) == mid:
                out.append("%s%de%d" % (rng.choice(["", "-"]), m, p))
                traps += 1
                break
    out = [t for t in out if abs(float(t)) < 3.4e38]  # no overflow (an error for both readers)
    out += ["16777217", "16777219", "33554435", "-16777217.0", "000123.4500", "+.5", "-0.0", "0e10", "-0",
            "1e-46", "1.4e-45", "-7e-46", "1.17549435e-38", "3.4028235e38", "340282356779733661637539395458142568447",
            "0.000000000000000000000000000000000000000000001", "9007199254740993", "9007199254740993e-10"]
    return out


def test_reader_rounds_every_decimal_form_as_an_istream(tmp_path):
    """The library's decimal parser (an exact fast path with strtof for the
    rest) against the oracle's istream reader on 200,000 numbers."""
    rng = np.random.default_rng(5)
    toks = _decimal_tokens(rng, 200000)
    toks += ["0"] * (-len(toks) % 8)
    rows = np.asarray(toks, dtype=object).reshape(-1, 8)
    lines = ["%d" % len(rows)] + [" ".join(list(r[:3]) + ["2"] + list(r[3:])) for r in rows]
    path = tmp_path / "numbers.txt"
    path.write_text("\n".join(lines) + "\n")
    got = upf.read_restart(str(path))
    cnt, orc = ol.read_restart(str(path), cap=len(rows))
    assert cnt == len(rows) and len(orc["type"]) == len(rows), "the istream stopped early"
    lib = np.concatenate([got["pos"][:, :3], got["vel"][:, :3], got["corr_pressure"][:, None],
                          got["predicted_density"][:, None]], 1)
    ref = np.concatenate([orc["pos"], orc["vel"], orc["corr_pressure"][:, None], orc["predicted_density"][:, None]], 1)
    bad = np.flatnonzero(lib.view(np.uint32).reshape(-1) != ref.view(np.uint32).reshape(-1))
    assert bad.size == 0, [toks[i] for i in bad[:5]]


def _dump_by_numpy(p, pos_sorted, key0_sorted):
    """DumpGPUParamsToFile's two records rebuilt with numpy from sorted
    positions: w = the uploaded z-index as float32; a particle's block is
    (int)(x * scales * (1 / block_size)) per axis in float32, row-major."""
    n = len(pos_sorted)
    rec = np.zeros((n, 4), np.float32)
    rec[:, :3] = pos_sorted[:, :3]
    rec[:, 3] = key0_sorted.astype(np.float32)
    dim = 1024 // p.block_size
    inv = np.float32(1.0) / np.float32(p.block_size)
    b = [(pos_sorted[:, a].astype(np.float32) * np.float32(p.scales[a]) * inv).astype(np.int64) for a in range(3)]
    lin = (dim * dim) * b[2] + dim * b[1] + b[0]
    u, first, cnt = np.unique(lin, return_index=True, return_counts=True)
    blocks = np.stack([u, first, cnt], 1).astype(np.uint32)
    return (np.uint32([n]).tobytes() + rec.tobytes(), np.uint32([len(u)]).tobytes() + blocks.tobytes())


def _morton_keys(p, pos):
    c = [np.floor(pos[:, a].astype(np.float32) * np.float32(p.scales[a])).astype(np.uint64) for a in range(3)]
    key = np.zeros(len(pos), np.uint64)
    for bit in range(10):
        for a in range(3):
            key |= ((c[a] >> np.uint64(bit)) & np.uint64(1)) << np.uint64(3 * bit + a)
    return key


@pytest.mark.parametrize("steps", [0, 2])
def test_oracle_cell_dump_equals_numpy_restatement(tmp_path, steps):
    """The oracle's binary dump against a numpy restatement from its sorted
    state; two dumps append two records."""
    import scenes
    p, pos, typ, nf = scenes.dam_break(8)
    o = ol.OracleSim(p, pos, None, typ, omp=False)
    o.step(steps, upf.UPF_PCISPH_STATIC) if steps else None
    pp, bp = str(tmp_path / "particles.bin"), str(tmp_path / "blocks.bin")
    o.dump_cells(pp, bp)
    st = o.slot_state()
    key0 = _morton_keys(p, np.asarray(pos))[st["ids"]]
    exp_p, exp_b = _dump_by_numpy(p, st["pos"], key0)
    assert read_bytes(pp) == exp_p and read_bytes(bp) == exp_b
    o.dump_cells(pp, bp)
    recs = upf.read_cell_dump(pp, bp)
    assert len(recs) == 2
    parts, blocks = recs[1]
    assert parts.tobytes() == exp_p[4:] and blocks.tobytes() == exp_b[4:]
    # the blocks tile the sorted particles: contiguous ranges covering every slot once
    order = np.argsort(blocks[:, 1])
    ends = blocks[order, 1] + blocks[order, 2]
    assert blocks[order[0], 1] == 0 and np.array_equal(ends[:-1], blocks[order[1:], 1]) and ends[-1] == len(typ)
