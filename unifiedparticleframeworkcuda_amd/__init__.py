"""MI355X-native per-step particle pipeline (PCISPH / WCSPH / rigid-fluid
coupling) of the Unified Particle Framework, behind a C ABI (include/upf.h).

This module is the host-side mirror of the reference's simulation-system API
(UnifiedPhysics + UnifiedConstants + Scene, src/cpu/UnifiedPhysics.h,
src/sceneConfigure/UnifiedConstants.h, src/sceneConfigure/Scene.h) for
Python callers: it loads the in-tree HIP library ``libupf.so`` through
ctypes and raises when the library is missing -- there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

from ._abi import (Params, DamBreakDesc, RigidSceneDesc, Body, Tuning, ptr, ERROR_NAMES,  # noqa: F401
                   UPF_OK, UPF_ERR_RANGE, UPF_LIQUID, UPF_FROZEN, UPF_RIGID, UPF_PCISPH_STATIC,
                   UPF_WCSPH_STATIC, UPF_PCISPH_RIGID, UPF_SPH_STATIC, UPF_BOUNDARY_CPU_REF, UPF_BOUNDARY_AKINCI,
                   UPF_TEXT_LIQUID, UPF_TEXT_RIGID, UPF_TEXT_POSITIONS, UPF_TEXT_RESTART)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libupf.so")

_lib = None


class UPFError(RuntimeError):
    def __init__(self, fn, code):
        super().__init__("%s failed: %s (%d)" % (fn, ERROR_NAMES.get(code, "?"), code))
        self.code = code


def lib():
    """The HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libupf.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        pf = C.POINTER(C.c_float)
        pi = C.POINTER(C.c_int)
        pu = C.POINTER(C.c_uint32)
        pp = C.POINTER(Params)
        vp = C.c_void_p
        L.upf_params_init.argtypes = [C.c_float, pp]
        L.upf_params_set_domain.argtypes = [pp, pf, pf, pf]
        L.upf_build_luts.argtypes = [pp, pf]
        L.upf_compute_density_error_factor.argtypes = [pp, pi]
        L.upf_scene_dam_break.argtypes = [C.POINTER(DamBreakDesc), pp, pi, pi, pf, pi]
        L.upf_scene_rigid.argtypes = [C.POINTER(RigidSceneDesc), pp, pi, pi, pi, pf, pi, pi, pf, C.POINTER(Body)]
        L.upf_create.argtypes = [pp, C.c_int, C.POINTER(vp)]
        L.upf_destroy.argtypes = [vp]
        L.upf_set_params.argtypes = [vp, pp]
        L.upf_upload_particles.argtypes = [vp, C.c_int, pf, pf, pi]
        L.upf_step.argtypes = [vp, C.c_int, C.c_int]
        L.upf_download.argtypes = [vp, pf, pf, pf, pf, pf, pf, pi]
        L.upf_get_step_stats.argtypes = [vp, C.c_int, pi, pf]
        L.upf_sort_and_bin.argtypes = [vp]
        L.upf_download_sorted.argtypes = [vp, pu, pi, pu, pi]
        L.upf_download_neighbors.argtypes = [vp, pi, pi, C.c_int]
        L.upf_keys.argtypes = [pp, C.c_int, pf, pu]
        L.upf_radix_sort.argtypes = [C.c_int, pu, pu, pi]
        L.upf_debug_box_cells.argtypes = [pp, C.c_float, C.c_float, C.c_float, pu, C.c_int]
        L.upf_version.restype = C.c_char_p
        L.upf_batch_create.argtypes = [C.c_int, pp, pi, C.c_int, C.POINTER(vp)]
        L.upf_batch_destroy.argtypes = [vp]
        L.upf_batch_upload.argtypes = [vp, pf, pf, pi]
        L.upf_batch_step.argtypes = [vp, C.c_int, C.c_int]
        L.upf_batch_download.argtypes = [vp, pf, pf, pf]
        L.upf_batch_get_stats.argtypes = [vp, C.c_int, pi, pf]
        L.upf_batch_sim.argtypes = [vp, C.POINTER(vp)]
        L.upf_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
        L.upf_get_tuning.argtypes = [vp, C.POINTER(Tuning)]
        L.upf_profile_enable.argtypes = [vp, C.c_int]
        L.upf_profile_reset.argtypes = [vp]
        L.upf_profile_read.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_longlong)]
        L.upf_profile_name.restype = C.c_char_p
        L.upf_obj_read.argtypes = [C.c_char_p, pf, C.c_int, pi]
        L.upf_mesh_fill.argtypes = [pf, C.c_int, pf, C.c_float, pf, C.c_int, pi]
        L.upf_save_text.argtypes = [vp, C.c_char_p, C.c_int]
        L.upf_write_text.argtypes = [C.c_char_p, C.c_int, C.c_int, pf, pi, pf, pf, pf]
        L.upf_read_restart.argtypes = [C.c_char_p, C.c_int, pi, pf, pf, pi, pf, pf]
        L.upf_upload_restart.argtypes = [vp, C.c_int, pf, pf, pi, pf, pf]
        L.upf_dump_cells.argtypes = [vp, C.c_char_p, C.c_char_p]
        L.upf_bodies_from_particles.argtypes = [pp, C.c_int, pf, pf, pf, pi, pf, C.POINTER(Body), C.c_int, pi]
        for name, args in (("upf_upload_bodies", [vp, C.c_int, C.POINTER(Body), pi, pf]),
                           ("upf_download_bodies", [vp, C.POINTER(Body)])):
            if hasattr(L, name):
                getattr(L, name).argtypes = args
        _lib = L
    return _lib


def check(fn, rc):
    if rc != UPF_OK:
        raise UPFError(fn, rc)
    return rc


# ---------------------------------------------------------------- set-up
def make_params(spacing=0.01, dt=None):
    """UnifiedConstants for a particle spacing (src/sceneConfigure/UnifiedConstants.cpp:7-191)."""
    p = Params()
    check("upf_params_init", lib().upf_params_init(spacing, C.byref(p)))
    if dt is not None:
        p.dt = dt
    return p


def load_obj(path):
    """Triangles of a Wavefront OBJ file as the reference's ObjLoader reads
    them (src/cpu/ObjLoader.cpp:18-258): (n, 3, 3) float32, file order."""
    L = lib()
    n = C.c_int()
    check("upf_obj_read", L.upf_obj_read(os.fsencode(path), None, 0, C.byref(n)))
    out = np.zeros((n.value, 3, 3), np.float32)
    if n.value:
        check("upf_obj_read", L.upf_obj_read(os.fsencode(path), ptr(out), n.value, C.byref(n)))
    return out


def mesh_fill(faces, scale, spacing):
    """ObjectLoader::LoadParticles (src/cpu/ObjectLoader.cpp:8-97) on the GPU:
    the lattice points inside the triangles `faces` ((n, 3, 3)) or on them,
    (point - box minimum) * scale, in the reference's order; (m, 3) float32."""
    L = lib()
    f = np.ascontiguousarray(faces, np.float32).reshape(-1, 9)
    s3 = np.broadcast_to(np.asarray(scale, np.float32), (3,)).copy()
    sc = (C.c_float * 3)(*[float(v) for v in s3])
    n = C.c_int()
    if len(f) == 0:
        return np.zeros((0, 3), np.float32)
    # the lattice holds at most as many particles as it has points: the box
    # of the faces (from the origin) in steps of spacing / scale, counted as
    # the reference's loops count them
    v = f.reshape(-1, 3)
    steps = (np.maximum(v.max(0), 0).astype(np.float32) - np.minimum(v.min(0), 0).astype(np.float32)) * s3 \
        / np.float32(spacing)
    cap = int(np.prod(np.floor(np.maximum(steps, -1)) + 1)) if np.all(steps >= 0) else 0
    out = np.zeros((max(cap, 1), 4), np.float32)
    rc = L.upf_mesh_fill(ptr(f), len(f), sc, float(spacing), ptr(out), cap, C.byref(n))
    if rc == UPF_ERR_RANGE:  # not expected: count, then fill
        out = np.zeros((n.value, 4), np.float32)
        rc = L.upf_mesh_fill(ptr(f), len(f), sc, float(spacing), ptr(out), n.value, C.byref(n))
    check("upf_mesh_fill", rc)
    return out[:n.value, :3].copy()


def model_particles(path, scale, spacing, offset=(0.0, 0.0, 0.0)):
    """Particle positions of an OBJ model placed as Scene::CreateFullParticleObject
    places them (src/sceneConfigure/Scene.cpp:797-824): the filled lattice plus
    the object's offset (float32 additions)."""
    pos = mesh_fill(load_obj(path), scale, spacing)
    return pos + np.asarray(offset, np.float32)


def write_text(path, which, pos4, types, vel4=None, corr_pressure=None, predicted_density=None):
    """The reference's particle files (src/cpu/UnifiedIO.cpp) from host arrays:
    which = UPF_TEXT_LIQUID / _RIGID / _POSITIONS, or UPF_TEXT_RESTART (then
    velocity, correction pressure and predicted density are written too)."""
    pos4 = np.ascontiguousarray(pos4, np.float32)
    types = np.ascontiguousarray(types, np.int32)
    n = len(types)
    arr = [None if a is None else np.ascontiguousarray(a, np.float32) for a in (vel4, corr_pressure,
                                                                              predicted_density)]
    check("upf_write_text", lib().upf_write_text(os.fsencode(path), int(which), n, ptr(pos4), ptr(types, C.c_int),
                                                 *[ptr(a) for a in arr]))


def read_restart(path):
    """ReadInFluidParticleNum + ReadInFluidParticlesPCISPH (src/cpu/UnifiedIO.cpp:168-220):
    dict of pos (n, 4), vel (n, 4), type, corr_pressure, predicted_density."""
    L = lib()
    n = C.c_int()
    check("upf_read_restart", L.upf_read_restart(os.fsencode(path), 0, C.byref(n), None, None, None, None, None))
    m = n.value
    out = dict(pos=np.zeros((m, 4), np.float32), vel=np.zeros((m, 4), np.float32), type=np.zeros(m, np.int32),
               corr_pressure=np.zeros(m, np.float32), predicted_density=np.zeros(m, np.float32))
    if m:
        check("upf_read_restart", L.upf_read_restart(os.fsencode(path), m, C.byref(n), ptr(out["pos"]),
                                                     ptr(out["vel"]), ptr(out["type"], C.c_int),
                                                     ptr(out["corr_pressure"]), ptr(out["predicted_density"])))
    return out


def read_cell_dump(particles_path, blocks_path):
    """The records of the binary dumps (DumpGPUParamsToFile): a list of
    (particles (n, 4) float32, blocks (m, 3) uint32 as (block index, first, count))."""
    pb = np.fromfile(particles_path, np.uint8)
    bb = np.fromfile(blocks_path, np.uint8)
    out, i, j = [], 0, 0
    while i < len(pb):
        n = int(pb[i:i + 4].view(np.uint32)[0])
        parts = pb[i + 4:i + 4 + 16 * n].view(np.float32).reshape(n, 4)
        i += 4 + 16 * n
        m = int(bb[j:j + 4].view(np.uint32)[0])
        blocks = bb[j + 4:j + 4 + 12 * m].view(np.uint32).reshape(m, 3)
        j += 4 + 12 * m
        out.append((parts, blocks))
    return out


def bodies_from_particles(p, pos4, mass, vel4=None):
    """Rigid bodies of solid particles as the reference's scene set-up makes
    them (FreezeParticle + AddToRigidBody, src/cpu/UnifiedPhysics.cpp:825-883,
    1657-1847): (body_id per particle (-1: none), rel4 (position - centre of
    mass, w = mass), list of Body)."""
    L = lib()
    pos4 = np.ascontiguousarray(pos4, np.float32)
    m = np.ascontiguousarray(np.broadcast_to(np.asarray(mass, np.float32), (len(pos4),)), np.float32)
    v = None if vel4 is None else np.ascontiguousarray(vel4, np.float32)
    n = len(pos4)
    bid = np.zeros(n, np.int32)
    rel = np.zeros((n, 4), np.float32)
    nb = C.c_int()
    check("upf_bodies_from_particles", L.upf_bodies_from_particles(C.byref(p), n, ptr(pos4), ptr(v), ptr(m),
                                                                   ptr(bid, C.c_int), None, None, 0, C.byref(nb)))
    arr = (Body * max(nb.value, 1))()
    check("upf_bodies_from_particles", L.upf_bodies_from_particles(C.byref(p), n, ptr(pos4), ptr(v), ptr(m),
                                                                   ptr(bid, C.c_int), ptr(rel), arr, nb.value,
                                                                   C.byref(nb)))
    return bid, rel, list(arr)[:nb.value]


def luts(p):
    out = np.zeros(5 * 301, np.float32)
    check("upf_build_luts", lib().upf_build_luts(C.byref(p), ptr(out)))
    return out.reshape(5, 301)


def density_error_factor(p):
    """Fills p.density_error_factor (ComputeDensityErrorFactor, src/cpu/UnifiedPhysics.cpp:5525)."""
    mx = C.c_int()
    check("upf_compute_density_error_factor", lib().upf_compute_density_error_factor(C.byref(p), C.byref(mx)))
    return p.density_error_factor, mx.value


def dam_break(p, nx, ny, nz, box, jitter_rel=0.001, wall_rel=1.0, lattice_rel=1.0025, seed=0):
    """Seeded dam-break scene; sets the domain of p.  Returns (pos4, type, n_fluid)."""
    d = DamBreakDesc(nx=nx, ny=ny, nz=nz, jitter_rel=jitter_rel, wall_rel=wall_rel, lattice_rel=lattice_rel,
                     seed=seed)
    d.box[:] = list(box)
    nf, nw = C.c_int(), C.c_int()
    check("upf_scene_dam_break", lib().upf_scene_dam_break(C.byref(d), C.byref(p), C.byref(nf), C.byref(nw),
                                                             None, None))
    n = nf.value + nw.value
    pos = np.zeros((n, 4), np.float32)
    typ = np.zeros(n, np.int32)
    check("upf_scene_dam_break", lib().upf_scene_dam_break(C.byref(d), C.byref(p), C.byref(nf), C.byref(nw),
                                                             ptr(pos), ptr(typ, C.c_int)))
    return pos, typ, nf.value


def standard_dam_break(nx, ny, nz, spacing=0.01, dt=0.001, seed=0, box_per_cell=(0.02, 0.01, 0.01)):
    """The dam break of the benchmark configurations: an nx*ny*nz lattice in
    the corner of a box of (2 nx, ny, nz) spacings, dt = 1 ms."""
    p = make_params(spacing, dt)
    box = (box_per_cell[0] * nx * spacing / 0.01, box_per_cell[1] * ny * spacing / 0.01,
           box_per_cell[2] * nz * spacing / 0.01)
    pos, typ, nf = dam_break(p, nx, ny, nz, box, seed=seed)
    density_error_factor(p)
    return p, pos, typ, nf


def rigid_scene(p, nx, ny, nz, box, n_bodies, size=(0.05, 0.15), density=(300.0, 2000.0), jitter_rel=0.001,
                surface_rel=0.5, wall_rel=1.0, lattice_rel=1.0025, seed=2):
    """Fluid + walls + surface-sampled rigid boxes/spheres; sets the domain of p.
    Returns (pos4, type, body_id, rel4, bodies, counts)."""
    d = RigidSceneDesc(nx=nx, ny=ny, nz=nz, n_bodies=n_bodies, size_min=size[0], size_max=size[1],
                       density_min=density[0], density_max=density[1], jitter_rel=jitter_rel,
                       surface_rel=surface_rel, wall_rel=wall_rel, lattice_rel=lattice_rel, seed=seed)
    d.box[:] = list(box)
    nf, nr, nw = C.c_int(), C.c_int(), C.c_int()
    L = lib()
    check("upf_scene_rigid", L.upf_scene_rigid(C.byref(d), C.byref(p), C.byref(nf), C.byref(nr), C.byref(nw),
                                               None, None, None, None, None))
    n = nf.value + nr.value + nw.value
    pos = np.zeros((n, 4), np.float32)
    typ = np.zeros(n, np.int32)
    bid = np.zeros(n, np.int32)
    rel = np.zeros((n, 4), np.float32)
    bodies = (Body * max(n_bodies, 1))()
    check("upf_scene_rigid", L.upf_scene_rigid(C.byref(d), C.byref(p), C.byref(nf), C.byref(nr), C.byref(nw),
                                               ptr(pos), ptr(typ, C.c_int), ptr(bid, C.c_int), ptr(rel), bodies))
    return pos, typ, bid, rel, list(bodies)[:n_bodies], (nf.value, nr.value, nw.value)


def standard_rigid_scene(nx=64, ny=64, nz=128, n_bodies=64, spacing=0.01, dt=0.001, seed=2,
                         mode=UPF_BOUNDARY_AKINCI):
    """The rigid-fluid coupling workload: an nx*ny*nz fluid block (x, y up, z) in a
    container 2 nx spacings long, nz deep, tall enough for the bodies stacked in
    layers 0.2 m apart above the fluid, from where they drop onto it;
    Akinci boundary volumes, surface tension 0.05, adhesion 0.01."""
    p = make_params(spacing, dt)
    lx, lz = 2 * nx * spacing, nz * spacing
    per_layer = max(1, int(lx * 1.0025 / 0.2)) * max(1, int(lz * 1.0025 / 0.2))
    layers = -(-n_bodies // per_layer)
    box = (lx, ny * spacing + 0.15 + 0.2 * layers + 0.05, lz)
    out = rigid_scene(p, nx, ny, nz, box, n_bodies, seed=seed)
    p.surface_tension_gamma = 0.05
    p.surface_adhesion_beta = 0.01
    p.boundary_pressure_mode = mode
    density_error_factor(p)
    return (p,) + out


# ------------------------------------------------------------ simulation
class Simulation:
    """One simulation on the current HIP device (the reference's UnifiedPhysics
    GPU path: InitCudaVariables + CudaParticlePhysics, src/cpu/UnifiedPhysics.cpp:615-643,4051-4141)."""

    def __init__(self, params, capacity):
        self.p = params.copy()
        self.h = C.c_void_p()
        check("upf_create", lib().upf_create(C.byref(self.p), int(capacity), C.byref(self.h)))
        self.n = 0

    def set_params(self, params):
        self.p = params.copy()
        check("upf_set_params", lib().upf_set_params(self.h, C.byref(self.p)))

    def upload(self, pos4, types, vel4=None):
        pos4 = np.ascontiguousarray(pos4, np.float32)
        types = np.ascontiguousarray(types, np.int32)
        vel = None if vel4 is None else np.ascontiguousarray(vel4, np.float32)
        self.n = len(types)
        check("upf_upload_particles", lib().upf_upload_particles(self.h, self.n, ptr(pos4), ptr(vel),
                                                                 ptr(types, C.c_int)))

    def upload_restart(self, state):
        """Scene::readFluidObjectFromFile + upload (src/sceneConfigure/Scene.cpp:648-697):
        `state` is what read_restart returns (or a path to read)."""
        if not isinstance(state, dict):
            state = read_restart(state)
        pos4 = np.ascontiguousarray(state["pos"], np.float32)
        types = np.ascontiguousarray(state["type"], np.int32)
        vel = np.ascontiguousarray(state["vel"], np.float32)
        cp = np.ascontiguousarray(state["corr_pressure"], np.float32)
        rho = np.ascontiguousarray(state["predicted_density"], np.float32)
        self.n = len(types)
        check("upf_upload_restart", lib().upf_upload_restart(self.h, self.n, ptr(pos4), ptr(vel),
                                                             ptr(types, C.c_int), ptr(cp), ptr(rho)))

    def save_text(self, path, which=UPF_TEXT_RESTART):
        """The particles in slot order as one of the reference's files
        (UnifiedIO, src/cpu/UnifiedIO.cpp); UPF_TEXT_RESTART reloads with
        upload_restart and continues exactly."""
        check("upf_save_text", lib().upf_save_text(self.h, os.fsencode(path), int(which)))

    def dump_cells(self, particles_path, blocks_path):
        """DumpGPUParamsToFile (src/cpu/UnifiedPhysics.cpp:411-474): appends the
        sorted particles and the block table of the current state to the two
        binary files (read them with read_cell_dump)."""
        check("upf_dump_cells", lib().upf_dump_cells(self.h, os.fsencode(particles_path), os.fsencode(blocks_path)))

    def upload_bodies(self, bodies, body_id, rel4):
        arr = (Body * len(bodies))(*bodies)
        bid = np.ascontiguousarray(body_id, np.int32)
        rel = np.ascontiguousarray(rel4, np.float32)
        check("upf_upload_bodies", lib().upf_upload_bodies(self.h, len(bodies), arr, ptr(bid, C.c_int), ptr(rel)))

    def step(self, nsteps=1, method=UPF_PCISPH_STATIC):
        check("upf_step", lib().upf_step(self.h, int(nsteps), int(method)))

    def set_tuning(self, tuning):
        t = tuning if isinstance(tuning, Tuning) else Tuning.from_dict(tuning)
        check("upf_set_tuning", lib().upf_set_tuning(self.h, C.byref(t)))

    def tuning(self):
        t = Tuning()
        check("upf_get_tuning", lib().upf_get_tuning(self.h, C.byref(t)))
        return t

    def state(self):
        n = self.n
        out = dict(pos=np.zeros((n, 4), np.float32), vel=np.zeros((n, 4), np.float32),
                   rho=np.zeros(n, np.float32), p=np.zeros(n, np.float32),
                   force=np.zeros((n, 4), np.float32), psi=np.zeros(n, np.float32), type=np.zeros(n, np.int32))
        check("upf_download", lib().upf_download(self.h, ptr(out["pos"]), ptr(out["vel"]), ptr(out["rho"]),
                                                 ptr(out["p"]), ptr(out["force"]), ptr(out["psi"]),
                                                 ptr(out["type"], C.c_int)))
        return out

    def bodies(self, nb):
        arr = (Body * nb)()
        check("upf_download_bodies", lib().upf_download_bodies(self.h, arr))
        return list(arr)

    def stats(self, max_steps=1 << 20):
        K = np.zeros(max_steps, np.int32)
        r = np.zeros(max_steps, np.float32)
        m = lib().upf_get_step_stats(self.h, max_steps, ptr(K, C.c_int), ptr(r))
        return K[:m], r[:m]

    def sort_and_bin(self):
        check("upf_sort_and_bin", lib().upf_sort_and_bin(self.h))

    def neighbors(self, cap=None):
        """(counts per sorted slot, concatenated lists) of the current state,
        after sorting it (non-liquid particles have empty lists)."""
        n = self.n
        cap = cap or n * self.p.max_neighbors
        counts = np.zeros(n, np.int32)
        flat = np.zeros(cap, np.int32)
        w = lib().upf_download_neighbors(self.h, ptr(counts, C.c_int), ptr(flat, C.c_int), cap)
        if w < 0:
            raise UPFError("upf_download_neighbors", -w)
        return counts, flat[:w]

    def sorted_state(self):
        n = self.n
        keys = np.zeros(n, np.uint32)
        ids = np.zeros(n, np.int32)
        nc = C.c_int()
        check("upf_download_sorted", lib().upf_download_sorted(self.h, ptr(keys, C.c_uint32), ptr(ids, C.c_int),
                                                               None, C.byref(nc)))
        cells = np.zeros((nc.value, 3), np.uint32)
        check("upf_download_sorted", lib().upf_download_sorted(self.h, None, None, ptr(cells, C.c_uint32),
                                                               C.byref(nc)))
        return keys, ids, cells

    def close(self):
        if self.h:
            lib().upf_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_keys(p, pos4):
    pos4 = np.ascontiguousarray(pos4, np.float32)
    keys = np.zeros(len(pos4), np.uint32)
    check("upf_keys", lib().upf_keys(C.byref(p), len(pos4), ptr(pos4), ptr(keys, C.c_uint32)))
    return keys


def device_radix_sort(keys):
    keys = np.ascontiguousarray(keys, np.uint32)
    out = np.zeros_like(keys)
    perm = np.zeros(len(keys), np.int32)
    check("upf_radix_sort", lib().upf_radix_sort(len(keys), ptr(keys, C.c_uint32), ptr(out, C.c_uint32),
                                                 ptr(perm, C.c_int)))
    return out, perm


def profile(handle):
    """Per-kernel-class device time {name: (total_ms, launches)} of a simulation handle."""
    L = lib()
    out = {}
    for c in range(L.upf_profile_classes()):
        ms, cnt = C.c_double(), C.c_longlong()
        L.upf_profile_read(handle, c, C.byref(ms), C.byref(cnt))
        if cnt.value:
            out[L.upf_profile_name(c).decode()] = (ms.value, cnt.value)
    return out


# ------------------------------------------------------------ batched form
def batch_dam_breaks(n_scenes=256, seed=0, count_range=(2048, 65536), aspect_range=(0.5, 2.0), spacing=0.01,
                     dt=0.001):
    """Uneven independent dam breaks: scene k has a fluid block of about c_k
    particles, c_k log-uniform in count_range, with block aspect ratio
    a_k = width / height uniform in aspect_range (nx = nz = a ny), in the
    corner of its own box of (2 nx, ny, nz) spacings with walls, jitter seed
    seed + k.  Returns a list of (params, pos4, type, n_fluid)."""
    rng = np.random.RandomState(seed)
    lo, hi = np.log(count_range[0]), np.log(count_range[1])
    scenes = []
    for k in range(n_scenes):
        c = float(np.exp(rng.uniform(lo, hi)))
        a = float(rng.uniform(*aspect_range))
        ny = max(2, int(round((c / (a * a)) ** (1.0 / 3.0))))
        nx = max(2, int(round(a * ny)))
        scenes.append(standard_dam_break(nx, ny, nx, spacing=spacing, dt=dt, seed=seed + k))
    return scenes


class Batch:
    """Many independent scenes stepped together (upf_batch_*): one launch per
    stage for all scenes, the whole step replayed as a hipGraph."""

    def __init__(self, scene_params, counts, max_record_steps=1024):
        arr = (Params * len(scene_params))(*scene_params)
        cnt = np.ascontiguousarray(counts, np.int32)
        self.n_scenes = len(scene_params)
        self.counts = cnt
        self.h = C.c_void_p()
        check("upf_batch_create", lib().upf_batch_create(self.n_scenes, arr, ptr(cnt, C.c_int),
                                                         int(max_record_steps), C.byref(self.h)))
        self.n = int(cnt.sum())

    @classmethod
    def from_scenes(cls, scenes, max_record_steps=1024):
        b = cls([sc[0] for sc in scenes], [len(sc[2]) for sc in scenes], max_record_steps)
        b.upload(np.concatenate([sc[1] for sc in scenes]), np.concatenate([sc[2] for sc in scenes]))
        return b

    def upload(self, pos4, types, vel4=None):
        pos4 = np.ascontiguousarray(pos4, np.float32)
        types = np.ascontiguousarray(types, np.int32)
        vel = None if vel4 is None else np.ascontiguousarray(vel4, np.float32)
        check("upf_batch_upload", lib().upf_batch_upload(self.h, ptr(pos4), ptr(vel), ptr(types, C.c_int)))

    def step(self, nsteps=1, use_graph=True):
        """use_graph: True / 1 every step one graph replay (the whole correction
        loop on the device), 2 the split graphs (head, loop chunks while the
        device's decision asks for more, tail), False / 0 no graphs."""
        mode = (1 if use_graph else 0) if isinstance(use_graph, bool) else int(use_graph)
        check("upf_batch_step", lib().upf_batch_step(self.h, int(nsteps), mode))

    def state(self):
        n = self.n
        out = dict(pos=np.zeros((n, 4), np.float32), vel=np.zeros((n, 4), np.float32), rho=np.zeros(n, np.float32))
        check("upf_batch_download", lib().upf_batch_download(self.h, ptr(out["pos"]), ptr(out["vel"]),
                                                             ptr(out["rho"])))
        return out

    def stats(self, max_steps=4096):
        K = np.zeros((max_steps, self.n_scenes), np.int32)
        r = np.zeros((max_steps, self.n_scenes), np.float32)
        m = lib().upf_batch_get_stats(self.h, max_steps, ptr(K, C.c_int), ptr(r))
        if m < 0:
            raise UPFError("upf_batch_get_stats", -m)
        m = min(m, max_steps)
        return K[:m], r[:m]

    def sim_handle(self):
        s = C.c_void_p()
        check("upf_batch_sim", lib().upf_batch_sim(self.h, C.byref(s)))
        return s

    def set_tuning(self, tuning):
        """Launch shapes (before the first graph capture, which bakes them in)."""
        t = tuning if isinstance(tuning, Tuning) else Tuning.from_dict(tuning)
        check("upf_set_tuning", lib().upf_set_tuning(self.sim_handle(), C.byref(t)))

    def close(self):
        if self.h:
            lib().upf_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
