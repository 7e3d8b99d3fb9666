"""ctypes bindings of the CPU oracle (oracle/liborc*.so) and of the reference
translation units compiled into oracle/_ref/libupfref.so.

Test infrastructure only: the oracle is the checker, never the thing measured.
"""
import contextlib
import ctypes as C
import os

import numpy as np

from unifiedparticleframeworkcuda_amd._abi import Params, Body, ptr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")

_libs = {}


def _bind(lib):
    f = C.c_float
    pf = C.POINTER(C.c_float)
    pi = C.POINTER(C.c_int)
    pu = C.POINTER(C.c_uint32)
    pp = C.POINTER(Params)
    lib.orc_params_init.argtypes = [f, pp]
    lib.orc_params_set_domain.argtypes = [pp, pf, pf, pf]
    lib.orc_density_error_factor.argtypes = [pp, pi]
    lib.orc_luts.argtypes = [f, pf, pf]
    lib.orc_lut_lookup.argtypes = [f, C.c_int, pf, pi, pf, pf, pf]
    lib.orc_ztable.argtypes = [pu]
    lib.orc_keys.argtypes = [pp, C.c_int, pf, pu]
    lib.orc_box_query.argtypes = [C.c_uint, pu, pu, pu, C.c_int]
    lib.orc_stable_sort.argtypes = [C.c_int, pu, pu, pi]
    lib.orc_sim_create.restype = C.c_void_p
    lib.orc_sim_create.argtypes = [pp, C.c_int, C.c_int, pf, pf, pi]
    lib.orc_sim_destroy.argtypes = [C.c_void_p]
    lib.orc_sim_upload_bodies.argtypes = [C.c_void_p, C.c_int, C.POINTER(Body), pi, pf, pf]
    lib.orc_sim_step.argtypes = [C.c_void_p, C.c_int, C.c_int]
    lib.orc_sim_download.argtypes = [C.c_void_p, pf, pf, pf, pf, pf, pf, pi, pu, pi, C.c_int]
    lib.orc_sim_download_bodies.argtypes = [C.c_void_p, C.POINTER(Body)]
    lib.orc_sim_stats.argtypes = [C.c_void_p, C.c_int, pi, pf]
    lib.orc_sim_neighbors.argtypes = [C.c_void_p, pi, pi, C.c_int]
    lib.orc_rigid_step.argtypes = [pf, C.c_int, pf, pf, pf, f, f, f, C.c_int]
    pd = C.POINTER(C.c_double)
    lib.orc_diff_run.argtypes = [pp, C.c_int, C.c_int, pd, pd, pi, C.c_int, C.c_int, pd, pd, pd, pd, pd, pd]
    lib.orc_diff_kernels.argtypes = [f, C.c_int, pd, pd, pd, pd, pd, pd]
    lib.orc_obj_read.argtypes = [C.c_char_p, pf, C.c_int, pi]
    lib.orc_mesh_fill.argtypes = [pf, C.c_int, pf, f, pf, C.c_int, pi]
    lib.orc_write_text.argtypes = [C.c_char_p, C.c_int, C.c_int, pf, pi, pf, pf, pf]
    lib.orc_read_restart.argtypes = [C.c_char_p, C.c_int, pi, pf, pf, pi, pf, pf]
    lib.orc_sim_set_restart.argtypes = [C.c_void_p, pf, pf]
    lib.orc_dump_cells.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p]
    lib.orc_wall_weight.argtypes = [pp, C.c_int, pf, pf, pf, pf]
    lib.orc_bodies_from_particles.argtypes = [pp, C.c_int, pf, pf, pf, pi, pf, C.POINTER(Body), pi]
    lib.orc_inverse3.argtypes = [pf, pf]
    lib.orc_body_inertia.argtypes = [C.c_int, pf, pf, pf, pf, pf, pf, pf, pf]
    return lib


def oracle(omp=False):
    """The oracle library; omp=True gives the OpenMP build (same results)."""
    name = "liborc_omp.so" if omp else "liborc.so"
    if name not in _libs:
        path = os.path.join(ORACLE_DIR, name)
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run __graft_entry__.build() (%s missing)" % path)
        _libs[name] = _bind(C.CDLL(path))
    return _libs[name]


@contextlib.contextmanager
def planted(path):
    """Inside the block, oracle() (both builds) returns the library at `path`:
    the oracle built with a planted mistake (-DORC_PLANT=k), for the tests that
    show their pins fail on it."""
    saved = dict(_libs)
    lib = _bind(C.CDLL(path))
    _libs["liborc.so"] = lib
    _libs["liborc_omp.so"] = lib
    try:
        yield lib
    finally:
        _libs.clear()
        _libs.update(saved)


def ref_built():
    """True when oracle/_ref/libupfref.so exists (without loading it)."""
    return os.path.exists(os.path.join(ORACLE_DIR, "_ref", "libupfref.so"))


def ref_lib():
    """oracle/_ref/libupfref.so or None when it was not built (no reference tree)."""
    path = os.path.join(ORACLE_DIR, "_ref", "libupfref.so")
    if not os.path.exists(path):
        return None
    if "ref" not in _libs:
        lib = C.CDLL(path)
        pf = C.POINTER(C.c_float)
        pu = C.POINTER(C.c_uint32)
        lib.ref_constants.argtypes = [C.c_float, pf]
        lib.ref_luts.argtypes = [C.c_float, pf, pf, C.c_int, pf, pf, pf, pf]
        lib.ref_kernels.argtypes = [C.c_float, C.c_int, pf, pf, pf, pf]
        lib.ref_ztable.argtypes = [pu]
        lib.ref_box_query.argtypes = [C.c_int, C.POINTER(C.c_ushort), C.POINTER(C.c_ushort), pu, C.c_int, pu]
        lib.ref_rigid_step.argtypes = [pf, C.c_int, pf, pf, pf, C.c_float, C.c_float, C.c_float]
        lib.ref_obj_particles.argtypes = [C.c_char_p, C.c_float, C.c_float, C.c_float, C.c_float, pf, C.c_int]
        lib.ref_inverse3.argtypes = [pf, pf]
        lib.ref_save_text.argtypes = [C.c_char_p, C.c_int, C.c_int, pf, C.POINTER(C.c_int), C.c_int]
        lib.ref_read_restart.argtypes = [C.c_char_p, C.c_int, pf, pf, C.POINTER(C.c_int), pf, pf]
        _libs["ref"] = lib
    return _libs["ref"]


def obj_read(path):
    """The oracle's OBJ reader: (n, 3, 3) float32 triangles (raises on error codes)."""
    L = oracle()
    n = C.c_int()
    rc = L.orc_obj_read(os.fsencode(path), None, 0, C.byref(n))
    if rc:
        raise ValueError("orc_obj_read: %d" % rc)
    out = np.zeros((n.value, 3, 3), np.float32)
    if n.value:
        L.orc_obj_read(os.fsencode(path), ptr(out), n.value, C.byref(n))
    return out


def mesh_fill(faces, scale, spacing):
    """The oracle's ObjectLoader::LoadParticles: (m, 3) float32 positions."""
    L = oracle()
    f = np.ascontiguousarray(faces, np.float32).reshape(-1, 9)
    sc = (C.c_float * 3)(*[float(v) for v in np.broadcast_to(np.asarray(scale, np.float32), (3,))])
    n = C.c_int()
    fp = ptr(f) if len(f) else None
    L.orc_mesh_fill(fp, len(f), sc, float(spacing), None, 0, C.byref(n))
    out = np.zeros((n.value, 3), np.float32)
    if n.value:
        L.orc_mesh_fill(fp, len(f), sc, float(spacing), ptr(out), n.value, C.byref(n))
    return out


def ref_obj_particles(path, scale, spacing):
    """The reference's ObjLoader + LoadParticles compiled into oracle/_ref
    (None without the reference tree): (m, 3) float32."""
    L = ref_lib()
    if L is None:
        return None
    sc = [float(v) for v in np.broadcast_to(np.asarray(scale, np.float32), (3,))]
    n = L.ref_obj_particles(os.fsencode(path), sc[0], sc[1], sc[2], float(spacing), None, 0)
    out = np.zeros((n, 3), np.float32)
    if n:
        L.ref_obj_particles(os.fsencode(path), sc[0], sc[1], sc[2], float(spacing), ptr(out), n)
    return out


def write_text(path, which, pos4, types, vel4=None, cp=None, rho=None):
    """The oracle's UnifiedIO writers (iostream formatting, as the reference)."""
    arr = [None if a is None else np.ascontiguousarray(a, np.float32) for a in (pos4, vel4, cp, rho)]
    t = np.ascontiguousarray(types, np.int32)
    rc = oracle().orc_write_text(os.fsencode(path), int(which), len(t), ptr(arr[0]), ptr(t, C.c_int),
                                 ptr(arr[1]), ptr(arr[2]), ptr(arr[3]))
    if rc:
        raise OSError("orc_write_text: %d" % rc)


def read_restart(path, cap=1 << 20):
    """The oracle's ReadInFluidParticlesPCISPH (operator>> while the stream is
    good): (count stated by the file, dict of the particles read completely)."""
    n = C.c_int()
    out = dict(pos=np.zeros((cap, 3), np.float32), vel=np.zeros((cap, 3), np.float32),
               type=np.zeros(cap, np.int32), corr_pressure=np.zeros(cap, np.float32),
               predicted_density=np.zeros(cap, np.float32))
    done = oracle().orc_read_restart(os.fsencode(path), cap, C.byref(n), ptr(out["pos"]), ptr(out["vel"]),
                                     ptr(out["type"], C.c_int), ptr(out["corr_pressure"]),
                                     ptr(out["predicted_density"]))
    if done < 0:
        raise OSError("orc_read_restart: cannot open %s" % path)
    return n.value, {k: v[:done] for k, v in out.items()}


def wall_weight(p, pos4):
    """The oracle's DistanceToWall and WallWeight (CPU-path and GPU-path table
    index) at the points pos4: (dist, w_cpu, w_gpu) float32 arrays."""
    x = np.ascontiguousarray(pos4, np.float32)
    n = len(x)
    d, wc, wg = (np.zeros(n, np.float32) for _ in range(3))
    oracle().orc_wall_weight(C.byref(p), n, ptr(x), ptr(d), ptr(wc), ptr(wg))
    return d, wc, wg


def bodies_from_particles(p, pos4, mass, vel4=None):
    """The oracle's bodies of solid particles: (body_id, rel4, list of Body)."""
    x = np.ascontiguousarray(pos4, np.float32)
    n = len(x)
    m = np.ascontiguousarray(np.broadcast_to(np.asarray(mass, np.float32), (n,)), np.float32)
    v = None if vel4 is None else np.ascontiguousarray(vel4, np.float32)
    bid = np.zeros(n, np.int32)
    rel = np.zeros((n, 4), np.float32)
    nb = C.c_int()
    L = oracle()
    L.orc_bodies_from_particles(C.byref(p), n, ptr(x), ptr(v), ptr(m), ptr(bid, C.c_int), None, None, C.byref(nb))
    arr = (Body * max(nb.value, 1))()
    L.orc_bodies_from_particles(C.byref(p), n, ptr(x), ptr(v), ptr(m), ptr(bid, C.c_int), ptr(rel), arr, C.byref(nb))
    return bid, rel, list(arr)[:nb.value]


def inverse3(m9, ref=False):
    """Matrix3x3::CalculateInverse: the oracle's, or the reference's own
    (oracle/_ref; None without it).  Returns (ok, 3x3 float32)."""
    L = ref_lib() if ref else oracle()
    if L is None:
        return None
    m = np.ascontiguousarray(m9, np.float32).reshape(9)
    out = np.zeros(9, np.float32)
    ok = (L.ref_inverse3 if ref else L.orc_inverse3)(ptr(m), ptr(out))
    return bool(ok), out.reshape(3, 3)


def body_accumulate(pos4, mass, vel4=None):
    """AddToRigidBody over all given particles in order: (inertia 3x3, com,
    vel, angular momentum, mass), float32."""
    x = np.ascontiguousarray(pos4, np.float32)
    n = len(x)
    m = np.ascontiguousarray(np.broadcast_to(np.asarray(mass, np.float32), (n,)), np.float32)
    v = None if vel4 is None else np.ascontiguousarray(vel4, np.float32)
    I, c, vv, L3, M = np.zeros(9, np.float32), np.zeros(3, np.float32), np.zeros(3, np.float32), np.zeros(3, np.float32), np.zeros(1, np.float32)
    oracle().orc_body_inertia(n, ptr(x), ptr(v), ptr(m), ptr(I), ptr(c), ptr(vv), ptr(L3), ptr(M))
    return I.reshape(3, 3), c, vv, L3, float(M[0])


REF_TEXT_FILES = {1: "FluidParticles/particles%04d.txt", 2: "RigidParticles/rigid1Particles%04d.txt",
                  3: "particles%04d.txt"}


def ref_save_text(workdir, which, pos4, types, count=None):
    """Runs the reference's UnifiedIO writer `which` (1 liquid, 2 rigid, 3
    positions) with `workdir` as working directory; returns the bytes it wrote,
    or None without the reference tree."""
    L = ref_lib()
    if L is None:
        return None
    out = os.path.join(workdir, "K:", "SimulationData", "output")
    for d in ("FluidParticles", "RigidParticles"):
        os.makedirs(os.path.join(out, d), exist_ok=True)
    p4 = np.ascontiguousarray(pos4, np.float32)
    t = np.ascontiguousarray(types, np.int32)
    if count is None:
        count = int(np.sum(t == 2)) if which == 1 else int(np.sum((t == 0) | (t == 6)))
    k = L.ref_save_text(os.fsencode(workdir), which, len(t), ptr(p4), ptr(t, C.c_int), count)
    if k < 0:
        raise OSError("ref_save_text failed")
    with open(os.path.join(out, REF_TEXT_FILES[which] % k), "rb") as f:
        return f.read()


def ref_read_restart(path, cap=1 << 20):
    """The reference's ReadInFluidParticleNum + ReadInFluidParticlesPCISPH (the
    file must exist): (count, dict of arrays of that count), None without the
    reference tree."""
    L = ref_lib()
    if L is None:
        return None
    out = dict(pos=np.zeros((cap, 3), np.float32), vel=np.zeros((cap, 3), np.float32),
               type=np.zeros(cap, np.int32), corr_pressure=np.zeros(cap, np.float32),
               predicted_density=np.zeros(cap, np.float32))
    n = L.ref_read_restart(os.fsencode(path), cap, ptr(out["pos"]), ptr(out["vel"]), ptr(out["type"], C.c_int),
                           ptr(out["corr_pressure"]), ptr(out["predicted_density"]))
    return n, {k: v[:max(n, 0)] for k, v in out.items()}


def params(spacing=0.01, vmax=None, cmin=None, cmax=None):
    p = Params()
    oracle().orc_params_init(spacing, C.byref(p))
    if vmax is not None:
        v = np.asarray(vmax, np.float32)
        a = np.asarray(cmin, np.float32)
        b = np.asarray(cmax, np.float32)
        rc = oracle().orc_params_set_domain(C.byref(p), ptr(v), ptr(a), ptr(b))
        if rc != 0:
            raise ValueError("orc_params_set_domain: %d" % rc)
    return p


class OracleSim:
    """Stateful oracle simulation (AoS, sorted like the reference's particles_)."""

    def __init__(self, p, pos4, vel4, types, precision=32, omp=True):
        self.lib = oracle(omp)
        self.n = len(types)
        self.pos4 = np.ascontiguousarray(pos4, np.float32)
        v = None if vel4 is None else np.ascontiguousarray(vel4, np.float32)
        t = np.ascontiguousarray(types, np.int32)
        self.h = self.lib.orc_sim_create(C.byref(p), precision, self.n, ptr(self.pos4), ptr(v), ptr(t, C.c_int))

    def upload_bodies(self, bodies, body_id, rel4, pmass=None):
        arr = (Body * len(bodies))(*bodies)
        bid = np.ascontiguousarray(body_id, np.int32)
        rel = np.ascontiguousarray(rel4, np.float32)
        pm = None if pmass is None else np.ascontiguousarray(pmass, np.float32)
        self.lib.orc_sim_upload_bodies(self.h, len(bodies), arr, ptr(bid, C.c_int), ptr(rel), ptr(pm))

    def step(self, nsteps, method):
        rc = self.lib.orc_sim_step(self.h, nsteps, method)
        if rc != 0:
            raise RuntimeError("oracle step failed: %d" % rc)

    def state(self, pcisph=True):
        n = self.n
        out = dict(pos=np.zeros((n, 4), np.float32), vel=np.zeros((n, 4), np.float32),
                   rho=np.zeros(n, np.float32), p=np.zeros(n, np.float32),
                   force=np.zeros((n, 4), np.float32), psi=np.zeros(n, np.float32),
                   type=np.zeros(n, np.int32), keys=np.zeros(n, np.uint32), ids=np.zeros(n, np.int32))
        self.lib.orc_sim_download(self.h, ptr(out["pos"]), ptr(out["vel"]), ptr(out["rho"]), ptr(out["p"]),
                                  ptr(out["force"]), ptr(out["psi"]), ptr(out["type"], C.c_int),
                                  ptr(out["keys"], C.c_uint32), ptr(out["ids"], C.c_int), 1 if pcisph else 0)
        return out

    def set_restart(self, corr_pressure, predicted_density):
        """Correction pressure and predicted density of a restart file, by upload index."""
        cp = np.ascontiguousarray(corr_pressure, np.float32)
        rho = np.ascontiguousarray(predicted_density, np.float32)
        self.lib.orc_sim_set_restart(self.h, ptr(cp), ptr(rho))

    def dump_cells(self, particles_path, blocks_path):
        """The oracle's DumpGPUParamsToFile records, appended to the two files."""
        rc = self.lib.orc_dump_cells(self.h, os.fsencode(particles_path), os.fsencode(blocks_path))
        if rc:
            raise OSError("orc_dump_cells: %d" % rc)

    def slot_state(self, pcisph=True, ids=None):
        """state() in the order of the oracle's particles (the CPU path's
        particles_, sorted at the end of every step), or in the order of the
        upload indices `ids`."""
        st = self.state(pcisph)
        ids = st["ids"] if ids is None else ids
        return {k: (v[ids] if k not in ("keys", "ids") else v) for k, v in st.items()}

    def bodies(self, nb):
        arr = (Body * nb)()
        self.lib.orc_sim_download_bodies(self.h, arr)
        return list(arr)

    def stats(self, max_steps=100000):
        K = np.zeros(max_steps, np.int32)
        r = np.zeros(max_steps, np.float32)
        m = self.lib.orc_sim_stats(self.h, max_steps, ptr(K, C.c_int), ptr(r))
        return K[:m], r[:m]

    def neighbors(self, cap=None):
        n = self.n
        cap = cap or n * 200
        counts = np.zeros(n, np.int32)
        flat = np.zeros(cap, np.int32)
        w = self.lib.orc_sim_neighbors(self.h, ptr(counts, C.c_int), ptr(flat, C.c_int), cap)
        assert w <= cap
        return counts, flat[:w]

    def __del__(self):
        try:
            self.lib.orc_sim_destroy(self.h)
        except Exception:
            pass


def _pd(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def diff_run(p, pos, vel, types, steps, K, gpos=None, gvel=None, precision=64, omp=True):
    """Differentiable PCISPH rollout on the oracle: returns (x_final, v_final)
    or, with gpos (dL/dx_final, n x 3), (x_final, v_final, dL/dx0, dL/dv0)."""
    n = len(types)
    x = np.ascontiguousarray(np.asarray(pos, np.float64)[:, :3])
    v = None if vel is None else np.ascontiguousarray(np.asarray(vel, np.float64)[:, :3])
    t = np.ascontiguousarray(types, np.int32)
    xo = np.zeros((n, 3))
    vo = np.zeros((n, 3))
    gp = None if gpos is None else np.ascontiguousarray(gpos, np.float64)
    gv = None if gvel is None else np.ascontiguousarray(gvel, np.float64)
    g0 = np.zeros((n, 3)) if gp is not None else None
    gv0 = np.zeros((n, 3)) if gp is not None else None
    rc = oracle(omp).orc_diff_run(C.byref(p), precision, n, _pd(x), _pd(v), ptr(t, C.c_int), steps, K, _pd(xo),
                                  _pd(vo), _pd(gp), _pd(gv), _pd(g0), _pd(gv0))
    if rc != 0:
        raise RuntimeError("orc_diff_run failed: %d" % rc)
    return (xo, vo) if gp is None else (xo, vo, g0, gv0)
