"""The C ABI library loads on a machine without a GPU and exports every
entry point include/upf.h declares; host-only helpers work without a device."""
import ctypes as C
import os
import re

import numpy as np

import unifiedparticleframeworkcuda_amd as upf
from unifiedparticleframeworkcuda_amd._abi import Params, DamBreakDesc, RigidSceneDesc, Body

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:int|const char\*|void)\s+(upf_\w+)\s*\(", text, re.M))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    lib = upf.lib()
    names = declared_functions()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_sizes_match_header_layout():
    # ctypes mirror vs the C layout: field count x 4 bytes plus alignment of the u64 seed
    n_arr = sum(1 for _, t in Params._fields_ if hasattr(t, "_length_"))
    assert C.sizeof(Params) == 4 * (len(Params._fields_) - n_arr) + 12 * n_arr
    assert C.sizeof(Body) % 4 == 0
    assert DamBreakDesc.seed.offset % 8 == 0 and RigidSceneDesc.seed.offset % 8 == 0


def test_version_string():
    assert b"gfx950" in upf.lib().upf_version()


def test_host_helpers_without_device():
    p = upf.make_params(0.01, 0.001)
    assert abs(p.support_radius - 0.0200002) < 1e-7
    pos, typ, nf = upf.dam_break(p, 8, 8, 8, (0.16, 0.08, 0.08))
    assert nf == 512 and np.sum(typ == upf.UPF_LIQUID) == 512
    assert p.block_size >= 2 and p.block_size >= max(p.scales) * p.support_radius
    f = typ == upf.UPF_LIQUID
    # fluid inside the collision box (up to the jitter), walls outside it
    jit = 0.001 * 0.01 + 1e-7
    assert np.all(pos[f, :3] >= np.array(p.collision_min) - jit)
    assert np.all(pos[f, :3] <= np.array(p.collision_max) + jit)
    w = ~f
    inside = np.all((pos[w, :3] > np.array(p.collision_min)) & (pos[w, :3] < np.array(p.collision_max)), axis=1)
    assert not inside.any()
    d, mx = upf.density_error_factor(p)
    assert d > 0 and mx == 32
    L = upf.luts(p)
    assert L.shape == (5, 301) and L[0, 300] == 0.0


def test_set_domain_rejects_cells_smaller_than_h():
    p = upf.make_params(0.01)
    # scales * h = 1024 * h / E just above a power of two -> block smaller than h
    E = 1024 * p.support_radius / 16.5
    vmax = (C.c_float * 3)(E, E, E)
    cmin = (C.c_float * 3)(0.0, 0.0, 0.0)
    cmax = (C.c_float * 3)(E * 0.5, E * 0.5, E * 0.5)
    rc = upf.lib().upf_params_set_domain(C.byref(p), vmax, cmin, cmax)
    assert rc == upf._abi.UPF_ERR_ARG


def test_cells_exactly_as_wide_as_h_are_rejected():
    """A cell exactly as wide as the scaled support radius lets a query box
    reach a fourth cell by rounding; set_domain and upf_create refuse it
    (upf_create checks parameters before it looks for a device)."""
    p = upf.make_params(0.01)
    E = 1024 * p.support_radius / 8.0  # scales * h = 8 = block size
    vmax = (C.c_float * 3)(E, E, E)
    cmin = (C.c_float * 3)(0.0, 0.0, 0.0)
    cmax = (C.c_float * 3)(E * 0.5, E * 0.5, E * 0.5)
    L = upf.lib()
    assert L.upf_params_set_domain(C.byref(p), vmax, cmin, cmax) == upf._abi.UPF_ERR_ARG
    q = upf.make_params(0.01)
    E = 1024 * q.support_radius / 7.8  # valid: scales * h = 7.8 < 8
    vmax = (C.c_float * 3)(E, E, E)
    assert L.upf_params_set_domain(C.byref(q), vmax, cmin, cmax) == upf._abi.UPF_OK
    q.scales[0] = q.block_size / q.support_radius
    h = C.c_void_p()
    assert L.upf_create(C.byref(q), 64, C.byref(h)) == upf._abi.UPF_ERR_ARG


def test_density_error_factor_rejects_reference_ub_spacing():
    p = upf.make_params(0.0125)
    rc = upf.lib().upf_compute_density_error_factor(C.byref(p), None)
    assert rc == upf._abi.UPF_ERR_ARG


def test_batch_create_validates_before_touching_a_device():
    """Scenes must share the physics (domains and density error factors may
    differ); counts must be positive.  Argument errors come before the device
    query, so they show without a GPU."""
    scenes = upf.batch_dam_breaks(3, seed=4, count_range=(200, 2000))
    ps = [s[0] for s in scenes]
    counts = [len(s[2]) for s in scenes]
    L = upf.lib()
    h = C.c_void_p()

    def create(params, cnt):
        arr = (Params * len(params))(*params)
        c = np.asarray(cnt, np.int32)
        return L.upf_batch_create(len(params), arr, c.ctypes.data_as(C.POINTER(C.c_int)), 16, C.byref(h))

    bad = [p.copy() for p in ps]
    bad[1].visc_const *= 2.0
    assert create(bad, counts) == upf._abi.UPF_ERR_ARG
    assert create(ps, [counts[0], 0, counts[2]]) == upf._abi.UPF_ERR_ARG
    rc = create(ps, counts)
    if rc == upf._abi.UPF_OK:  # a device is present
        L.upf_batch_destroy(h)
    else:
        assert rc == upf._abi.UPF_ERR_NO_DEVICE


def test_inconsistent_domains_are_rejected_before_the_device():
    """Cell tables hold grid_dim^3 entries indexed by key >> 3 block_shift:
    every creation call rejects params whose block_shift, block_size and
    grid_dim disagree (UPF_ERR_ARG, before any device call)."""
    L = upf.lib()
    base = upf.make_params(0.01, 0.001)
    upf.dam_break(base, 8, 8, 8, (0.16, 0.08, 0.08))
    assert base.block_size == 1 << base.block_shift and base.grid_dim == 1024 // base.block_size

    def variants():
        for field, value in (("block_shift", base.block_shift - 1), ("block_shift", base.block_shift + 1),
                             ("grid_dim", base.grid_dim * 2), ("grid_dim", base.grid_dim // 2),
                             ("block_size", base.block_size * 3 // 2), ("block_shift", 31)):
            q = base.copy()
            setattr(q, field, value)
            yield field, q

    h = C.c_void_p()
    for field, q in variants():
        assert L.upf_create(C.byref(q), 1000, C.byref(h)) == 2, field  # UPF_ERR_ARG
        arr = (Params * 2)(base, q)
        cnt = (C.c_int * 2)(100, 100)
        assert L.upf_batch_create(2, arr, cnt, 4, C.byref(h)) == 2, field  # UPF_ERR_ARG
        assert L.upf_adjoint_create(C.byref(q), 1000, 4, 2, 2, C.byref(h)) == 2, field  # UPF_ERR_ARG


def test_slot_counts_beyond_the_list_entry_range_are_rejected():
    """A packed list entry holds a slot in 23 bits: a simulation or batch of
    2^23 slots or more is rejected (UPF_ERR_ARG) before any device call."""
    L = upf.lib()
    p = upf.make_params(0.01, 0.001)
    upf.dam_break(p, 8, 8, 8, (0.16, 0.08, 0.08))
    h = C.c_void_p()
    assert L.upf_create(C.byref(p), 1 << 23, C.byref(h)) == 2  # UPF_ERR_ARG
    arr = (Params * 2)(p, p)
    cnt = (C.c_int * 2)(1 << 22, 1 << 22)
    assert L.upf_batch_create(2, arr, cnt, 4, C.byref(h)) == 2  # UPF_ERR_ARG
