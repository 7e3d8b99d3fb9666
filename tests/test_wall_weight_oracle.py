"""The wall weight function (Kulasegaram et al.; DistanceToWall / WallWeight,
src/cpu/UnifiedPhysics.cpp:5926-6060; the GPU path's DistanceToWallDevice /
WallWeightDevice, src/gpu/particles_kernel.cuh:612-693) in the oracle.

Pins: the wall distance against a brute-force minimum over the six faces of
the container; the weight against an explicit sum over the ten virtual wall
particles at their positions (two layers, one and two spacings aside on both
sides) with the table the reference's own WeightingKernel.cpp builds; its
place in the WCSPH density (added after the neighbour sum, at the current
position) and in the PCISPH predicted density (at the predicted position).
"""
import numpy as np
import pytest

import oracle_lib as ol
import scenes
import unifiedparticleframeworkcuda_amd as upf


def wall_params():
    p, pos, typ, nf = scenes.dam_break(8)
    p.add_wall_weight = 1
    return p, pos, typ, nf


def test_container_follows_the_collision_box():
    p, _, _, _ = wall_params()
    r = np.float32(p.particle_radius)
    for a in range(3):
        assert p.container_min[a] == np.float32(p.collision_min[a]) - r
        assert p.container_max[a] == np.float32(p.collision_max[a]) + r


def test_wall_distance_is_the_nearest_face(tmp_path):
    p, _, _, _ = wall_params()
    lo = np.float32(list(p.container_min))
    hi = np.float32(list(p.container_max))
    rng = np.random.default_rng(3)
    x = np.zeros((5000, 4), np.float32)
    x[:, :3] = lo - np.float32(0.05) + rng.random((5000, 3)).astype(np.float32) * (hi - lo + np.float32(0.1))
    d, _, _ = ol.wall_weight(p, x)
    faces = np.stack([x[:, 1] - lo[1], hi[1] - x[:, 1], x[:, 0] - lo[0], hi[0] - x[:, 0], x[:, 2] - lo[2],
                      hi[2] - x[:, 2]], 1).astype(np.float32)
    longest = np.max(hi - lo)
    exp = np.minimum(np.min(np.maximum(faces, np.float32(0)), 1), longest)
    assert np.array_equal(d, exp)


def test_wall_weight_is_the_sum_over_ten_virtual_wall_particles():
    p, _, _, _ = wall_params()
    h, r, s = np.float32(p.support_radius), np.float32(p.particle_radius), np.float32(p.spacing)
    lut = np.zeros(301, np.float32)
    upf_l = upf.luts(p)[0]  # M4 table (pinned against the reference's WeightingKernel.cpp)
    lut[:] = upf_l
    lo = np.float32(list(p.container_min))
    hi = np.float32(list(p.container_max))
    mid = (lo + hi) / np.float32(2)
    # points above the floor at every distance from 0 to past h - r
    ys = lo[1] + np.linspace(0, 1.1 * float(h), 2000).astype(np.float32)
    x = np.zeros((len(ys), 4), np.float32)
    x[:, 0], x[:, 1], x[:, 2] = mid[0], ys, mid[2]
    d, w_cpu, w_gpu = ol.wall_weight(p, x)
    assert np.array_equal(d, np.maximum(ys - lo[1], np.float32(0)))
    factor = np.float32(1.0 / float(h))
    m = np.float64(np.float32(p.initial_mass))
    for w, cpu in ((w_cpu, True), (w_gpu, False)):
        exp = np.zeros(len(d))
        for i, dist in enumerate(d):
            if dist >= h - r:
                continue
            d1 = np.float32(dist + r)
            d2 = np.float32(d1 + s)
            tot = 0.0
            # the ten virtual particles below the wall: (lateral offset, depth)
            for off, depth in [(0, d1), (0, d2)] + [(sgn * k * s, dep) for sgn in (-1, 1) for k in (1, 2)
                                                    for dep in (d1, d2)]:
                off = np.float32(off)
                dv = np.float32(np.sqrt(np.float32(depth * depth + off * off)))
                idx = int(np.float32(dv * factor) * np.float32(300)) if cpu else int(np.float32(dv * np.float32(300)) * factor)
                tot += 0.0 if idx >= 300 else float(lut[idx])
            exp[i] = m * tot
        assert np.allclose(w, exp, rtol=2e-6, atol=0), np.max(np.abs(w - exp))
    eff = d >= h - r
    assert np.all(w_cpu[eff] == 0) and np.all(w_cpu[~eff] > 0)
    assert np.all(np.diff(w_cpu) <= 0)  # non-increasing away from the wall


def _index_orders_differ(p, d):
    """Where the CPU path's table index (d * (1/h)) * 300 and the GPU path's
    (d * 300) * (1/h) differ for one of the ten virtual particles."""
    h, r, s = np.float32(p.support_radius), np.float32(p.particle_radius), np.float32(p.spacing)
    f = np.float32(1.0 / float(h))
    d1 = (d + r).astype(np.float32)
    d2 = (d1 + s).astype(np.float32)
    out = np.zeros(len(d), bool)
    for a, b in ((d1, 0), (d2, 0), (d1, s), (d1, 2 * s), (d2, s), (d2, 2 * s)):
        b = np.float32(b)
        dv = np.sqrt((a * a + b * b).astype(np.float32)).astype(np.float32)
        i1 = ((dv * f).astype(np.float32) * np.float32(300)).astype(np.float32).astype(np.int64)
        i2 = ((dv * np.float32(300)).astype(np.float32) * f).astype(np.float32).astype(np.int64)
        out |= (i1 != i2) & (np.minimum(i1, i2) < 300)
    return out & (d < h - r)


def _floor_points(p, ys):
    lo = np.float32(list(p.container_min))
    hi = np.float32(list(p.container_max))
    x = np.zeros((len(ys), 4), np.float32)
    x[:, 0], x[:, 1], x[:, 2] = (lo[0] + hi[0]) / 2, ys, (lo[2] + hi[2]) / 2
    return x


def test_cpu_and_gpu_table_index_orders_differ_exactly_where_they_should():
    p, _, _, _ = wall_params()
    lo1 = np.float32(p.container_min[1])
    ys = (lo1 + np.linspace(0, float(p.support_radius), 400000)).astype(np.float32)
    d, w_cpu, w_gpu = ol.wall_weight(p, _floor_points(p, ys))
    diff = _index_orders_differ(p, d)
    assert diff.sum() > 0
    assert np.array_equal(w_cpu != w_gpu, diff)


def _distinguishing_heights(p, count):
    lo1 = np.float32(p.container_min[1])
    ys = (lo1 + np.linspace(0, float(p.support_radius), 400000)).astype(np.float32)
    d, _, _ = ol.wall_weight(p, _floor_points(p, ys))
    return ys[_index_orders_differ(p, d)][:count]


def test_lone_particle_wcsph_density_uses_the_cpu_table_index():
    p, _, _, _ = wall_params()
    p.gas_const_wcsph = 1000.0
    p.dt = 0.000253
    y = _distinguishing_heights(p, 1)[0]
    x = _floor_points(p, np.float32([y]))
    o = ol.OracleSim(p, x, None, np.full(1, upf.UPF_LIQUID, np.int32), omp=False)
    o.step(1, upf.UPF_WCSPH_STATIC)
    _, w_cpu, w_gpu = ol.wall_weight(p, x)
    assert w_cpu[0] != w_gpu[0]
    m = np.float32(p.initial_mass)
    assert o.state(pcisph=False)["rho"][0] == np.float32(np.float32(np.float32(p.kernel_self) * m) + w_cpu[0])


def test_wcsph_density_adds_the_weight_after_the_neighbour_sum():
    p, pos, typ, nf = wall_params()
    p.gas_const_wcsph = 1000.0
    p.dt = 0.000253
    q = p.copy()
    q.add_wall_weight = 0
    a = ol.OracleSim(p, pos, None, typ, omp=False)
    b = ol.OracleSim(q, pos, None, typ, omp=False)
    a.step(1, upf.UPF_WCSPH_STATIC)
    b.step(1, upf.UPF_WCSPH_STATIC)
    ra, rb = a.state(pcisph=False)["rho"], b.state(pcisph=False)["rho"]
    _, w_cpu, _ = ol.wall_weight(p, pos)  # at the positions the density pass reads
    f = typ == upf.UPF_LIQUID
    assert np.array_equal(ra[f], (rb[f] + w_cpu[f]).astype(np.float32))
    assert np.any(w_cpu[f] > 0)


def test_pcisph_predicted_density_of_a_lone_particle_is_self_plus_wall():
    """A liquid particle at rest without neighbours near the floor (no
    gravity): every iteration's predicted density is m W(0) + the wall weight
    at its predicted position, which is its position."""
    p, _, _, _ = wall_params()
    p.gravity = 0.0  # at rest the predicted position is the position
    t = np.full(1, upf.UPF_LIQUID, np.int32)
    m = np.float32(p.initial_mass)
    told = 0
    # heights where the two table index orders give different weights (the
    # GPU path's must be used)
    for y in _distinguishing_heights(p, 200):
        o = ol.OracleSim(p, _floor_points(p, np.float32([y])), None, t, omp=False)
        o.step(1, upf.UPF_PCISPH_STATIC)
        st = o.state()
        _, w_cpu, w_gpu = ol.wall_weight(p, st["pos"])
        assert w_gpu[0] > 0
        assert st["rho"][0] == np.float32(np.float32(np.float32(p.kernel_self) * m) + w_gpu[0])
        told += w_cpu[0] != w_gpu[0]
    assert told > 0
