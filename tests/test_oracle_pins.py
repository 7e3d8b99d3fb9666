"""Pins of the oracle's step expressions by closed forms of a few particles.

The oracle restates the reference's CPU step (src/cpu/UnifiedPhysics.cpp);
these tests check what its per-pair and per-particle expressions compute
against values derived here in float64 from the published formulas:

* the kernels in closed form (cubic spline M4 = KernelM4, spiky gradient =
  kernelPressureGrad, viscosity Laplacian = kernelViscosityLaplacian, Akinci's
  cohesion and adhesion splines = kernelSplineSurfaceTension /
  kernelSplineSurfaceAdhesion, src/cpu/WeightingKernel.cpp:75-263), read the
  way the reference reads its 300-entry tables: the value at the start of bin
  floor(d * 300 / h);
* PCISPH (Solenthaler & Pajarola 2009, algorithm 1) as the CPU path runs it:
  external forces once (:5198-5250), then per iteration prediction with the
  collision clamp (:5426-5454, :5739-5757), predicted density and pressure
  accumulation (:5456-5491) and the corrective pressure force at the current
  positions (:5493-5522); symplectic Euler (:3705-3742) and the box clamp
  with damping -0.5 (:4000-4044);
* the Tait equation of state (:2491-2501) and Monaghan's symmetric pressure
  force with the viscosity term (:3009-3083);
* surface tension, adhesion, spring and damping pair forces (:1990-2030).

Every scene is a few particles placed so that each pair distance sits well
inside a table bin (the float32 and float64 bins agree).  The oracle can be
built with a planted mistake (-DORC_PLANT=k); test_planted_mistakes_fail_a_pin
builds each of the PLANTS below and shows that at least the pins named for it
fail, so the pins discriminate.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as ol
import unifiedparticleframeworkcuda_amd as upf

BOX = (0.6, 0.6, 0.6)
CMIN = (0.02, 0.02, 0.02)
CMAX = (0.58, 0.58, 0.58)


# ------------------------------------------------------------ kernels
def kernel(kind, r, h):
    """The reference's kernels in closed form (float64)."""
    s = r / h
    if kind == "w":  # cubic spline M4, 8 / (pi h^3) normalisation
        if r >= h:
            return 0.0
        f = 8.0 / (np.pi * h ** 3)
        return f * (1 - 6 * s * s + 6 * s ** 3) if s < 0.5 else f * 2 * (1 - s) ** 3
    if kind == "grad":  # spiky gradient / r
        return 0.0 if r == 0 or r > h else -45.0 / (np.pi * h ** 6) * (h - r) ** 2 / r
    if kind == "visc":
        return 0.0 if r > h else 45.0 / (np.pi * h ** 6) * (h - r)
    if kind == "tension":  # Akinci et al. 2013 cohesion spline
        c = 32.0 / (np.pi * h ** 9)
        if 0.5 * h < r <= h:
            return c * (h - r) ** 3 * r ** 3
        if 0 < r <= 0.5 * h:
            return c * (2 * (h - r) ** 3 * r ** 3 - h ** 6 / 64)
        return 0.0
    if kind == "adhesion":  # Akinci et al. 2013 adhesion spline
        if 0.5 * h < r <= h:
            return 0.007 / h ** 3.25 * (-4 * r * r / h + 6 * r - 2 * h) ** 0.25
        return 0.0
    raise ValueError(kind)


def table(kind, r, h):
    """The tabulated kernel: the value at the start of bin floor(r * 300 / h);
    r must lie well inside its bin."""
    u = r * 300.0 / h
    b = int(np.floor(u))
    assert 0.02 < u - b < 0.98, "pair distance %r too close to a table bin edge" % r
    return 0.0 if b >= 300 else kernel(kind, b * h / 300.0, h)


# ------------------------------------------------------------- scenes
def base_params(gravity=0.0, rho0=1000.0, K=None, mode=upf.UPF_BOUNDARY_CPU_REF):
    p = ol.params(0.01, BOX, CMIN, CMAX)
    p.gravity = gravity
    p.dt = 0.001
    p.rest_density = rho0
    p.density_error_factor = 2.0
    p.boundary_pressure_mode = mode
    if K is not None:
        p.min_loops = p.max_loops = K
    return p


def f64(x):
    return np.asarray(x, np.float32).astype(np.float64)


def run(p, pos, typ, method, vel=None, steps=1, bodies=None):
    pos4 = np.zeros((len(pos), 4), np.float32)
    pos4[:, :3] = pos
    v4 = None
    if vel is not None:
        v4 = np.zeros((len(pos), 4), np.float32)
        v4[:, :3] = vel
    o = ol.OracleSim(p, pos4, v4, np.asarray(typ, np.int32), precision=32, omp=False)
    if bodies is not None:
        o.upload_bodies(*bodies)
    o.step(steps, method)
    st = o.state(pcisph=method in (upf.UPF_PCISPH_STATIC, upf.UPF_PCISPH_RIGID))
    return o, st


def close(got, exp, rtol, what, atol=0.0):
    got = np.asarray(got, np.float64)
    exp = np.asarray(exp, np.float64)
    err = np.abs(got - exp).max()
    lim = rtol * np.abs(exp).max() + atol
    assert err <= lim, "%s: |got - expected| = %.3e > %.3e\n got %s\n exp %s" % (what, err, lim, got, exp)


def line(start, direction, gaps):
    e = np.asarray(direction, np.float64)
    e = e / np.linalg.norm(e)
    pts = [np.asarray(start, np.float64)]
    for g in gaps:
        pts.append(pts[-1] + g * e)
    return np.float32(pts)


# ------------------------------------------------- PCISPH model (float64)
def pcisph_model(p, x, v, K, liquid_mass=None):
    """K PCISPH iterations + symplectic Euler for liquid particles only (no
    walls, no bodies, zero viscosity contribution unless velocities differ)."""
    h = float(p.support_radius)
    m = float(p.initial_mass)
    rho0 = float(p.rest_density)
    dt = float(p.dt)
    delta = float(p.density_error_factor)
    mu = float(p.visc_const)
    g = float(p.gravity)
    n = len(x)
    lo, hi = f64(p.collision_min), f64(p.collision_max)
    nb = [[j for j in range(n) if j != i and np.linalg.norm(x[i] - x[j]) < h] for i in range(n)]
    # external forces: viscosity (rest volume m / rho0) and gravity
    F = np.zeros((n, 3))
    V0 = m / rho0
    for i in range(n):
        for j in nb[i]:
            F[i] -= (v[i] - v[j]) * V0 * V0 * mu * table("visc", np.linalg.norm(x[i] - x[j]), h)
        F[i, 1] -= m * g
    pc = np.zeros(n)
    Fp = np.zeros((n, 3))
    rho = np.zeros(n)
    for _ in range(K):
        vs = v + (F + Fp) * dt / m
        xs = x + vs * dt
        xs = np.clip(xs, lo, hi)
        for i in range(n):
            rho[i] = m * kernel("w", 0.0, h)
            for j in nb[i]:
                d = np.linalg.norm(xs[i] - xs[j])
                if d < h:
                    rho[i] += m * table("w", d, h)
        pc += delta * np.maximum(rho - rho0, 0.0)
        Fp = np.zeros((n, 3))
        for i in range(n):
            for j in nb[i]:
                d = np.linalg.norm(x[i] - x[j])
                Fp[i] -= m * m * (pc[i] / rho0 ** 2 + pc[j] / rho0 ** 2) * table("grad", d, h) * (x[i] - x[j])
    Ft = F + Fp
    v1 = v + Ft / m * dt
    x1 = x + v1 * dt
    return dict(rho=rho, p=pc, force=Ft, vel=v1, pos=x1)


def check_state(st, exp, x0, what):
    close(st["rho"], exp["rho"], 1e-6, what + " predicted density")
    close(st["p"], exp["p"], 1e-5, what + " correction pressure")
    close(st["force"][:, :3], exp["force"], 1e-5, what + " force")
    close(st["vel"][:, :3], exp["vel"], 1e-5, what + " velocity")
    # displacement, to the float32 resolution of the positions
    close(st["pos"][:, :3] - x0, exp["pos"] - x0, 1e-4, what + " displacement", atol=1e-7)


# ============================================================ the pins
def pin_pcisph_pressure_three_particles():
    """Three liquid particles on a line with unequal gaps (A-B 0.4517 h,
    B-C 0.7183 h, A-C beyond h), rest density 300 so that all are compressed,
    one and two correction iterations: densities m (W(0) + sum W), pressures
    delta (rho* - rho0) accumulated over the iterations, the symmetric
    corrective force -m^2 (p_i / rho0^2 + p_j / rho0^2) grad W pushing the
    particles apart, and the symplectic Euler update."""
    for K in (1, 2):
        p = base_params(rho0=300.0, K=K)
        h = float(p.support_radius)
        x = line((0.3, 0.3, 0.3), (3, 2, 6), (0.4517 * h, 0.7183 * h))
        o, st = run(p, x, [2, 2, 2], upf.UPF_PCISPH_STATIC)
        exp = pcisph_model(p, f64(x), np.zeros((3, 3)), K)
        check_state(st, exp, f64(x), "K=%d" % K)
        assert np.all(exp["p"] > 0) and len(set(np.round(exp["p"], 6))) == 3
        # A is pushed away from B, C away from B
        e = f64(x[1] - x[0])
        assert np.dot(st["force"][0, :3], e) < 0 < np.dot(st["force"][2, :3], e)


def pin_pcisph_viscosity_and_gravity():
    """Two liquid particles at 0.6217 h moving relative to each other, rest
    density 1000 (no pressure): the viscosity force -(v_i - v_j) V0^2 mu
    lap W with V0 = m / rho0 and gravity -m g; the relative velocity shrinks."""
    p = base_params(gravity=9.81)
    h = float(p.support_radius)
    x = line((0.3, 0.35, 0.3), (1, -2, 2), (0.6217 * h,))
    v = np.float32([[0.3, -0.1, 0.2], [-0.2, 0.25, 0.05]])
    o, st = run(p, x, [2, 2], upf.UPF_PCISPH_STATIC, vel=v)
    exp = pcisph_model(p, f64(x), f64(v), 3)
    assert np.all(exp["p"] == 0)
    check_state(st, exp, f64(x), "viscosity")
    dv0 = np.linalg.norm(f64(v[0] - v[1]))
    dv1 = np.linalg.norm(st["vel"][0, :3].astype(np.float64) - st["vel"][1, :3])
    assert dv1 < dv0


def pin_free_fall_and_floor_hit():
    """One liquid particle falling from rest under gravity: the symplectic
    Euler closed form y_n = y0 - g dt^2 n (n + 1) / 2, v_n = -g dt n until the
    step that crosses the floor, which puts it on the floor with velocity
    -0.5 v; then v = 0.5 g dt n_h - g dt k, y = floor + dt sum v."""
    p = base_params(gravity=9.81)
    g, dt = float(p.gravity), float(p.dt)
    floor = float(p.collision_min[1])
    y0 = floor + 0.004
    x = np.float32([[0.3, y0, 0.3]])
    y0 = float(x[0, 1])
    # step of the hit: first n with y0 - g dt^2 n (n + 1) / 2 < floor
    nh = next(n for n in range(1, 1000) if y0 - g * dt * dt * n * (n + 1) / 2 < floor)
    o = None
    for n in range(1, nh + 6):
        if o is None:
            pos4 = np.zeros((1, 4), np.float32)
            pos4[0, :3] = x[0]
            o = ol.OracleSim(p, pos4, None, np.int32([2]), precision=32, omp=False)
        o.step(1, upf.UPF_PCISPH_STATIC)
        st = o.state()
        if n < nh:
            ye, ve = y0 - g * dt * dt * n * (n + 1) / 2, -g * dt * n
        else:
            k = n - nh
            ve = 0.5 * g * dt * nh - g * dt * k
            ye = floor + dt * (0.5 * g * dt * nh * k - g * dt * k * (k + 1) / 2)
        assert abs(st["vel"][0, 1] - ve) <= 1e-5 * g * dt * nh, (n, st["vel"][0, 1], ve)
        assert abs(st["pos"][0, 1] - ye) <= 2e-7, (n, st["pos"][0, 1], ye)
        assert st["pos"][0, 0] == x[0, 0] and st["pos"][0, 2] == x[0, 2]
    assert nh > 5


def pin_predicted_position_clamp():
    """A liquid particle just above the floor moving down at 1 m/s, a lone
    boundary particle 0.3617 h below the floor (Akinci volume 0.7 / W(0)):
    its predicted position is clamped to the floor, so its predicted density
    is m W(0) + rho0 Psi W(0.3617 h), not the value at the unclamped
    prediction."""
    p = base_params(gravity=9.81, mode=upf.UPF_BOUNDARY_AKINCI)
    h = float(p.support_radius)
    floor = float(p.collision_min[1])
    x = np.float32([[0.3, floor + 0.0002, 0.3], [0.3, floor - 0.3617 * h, 0.3]])
    v = np.float32([[0.0, -1.0, 0.0], [0.0, 0.0, 0.0]])
    o, st = run(p, x, [2, 6], upf.UPF_PCISPH_STATIC, vel=v)
    m, rho0 = float(p.initial_mass), float(p.rest_density)
    w0 = kernel("w", 0.0, h)
    psi = 0.7 / w0
    d = float(floor) - float(x[1, 1])
    exp = m * w0 + rho0 * psi * table("w", d, h)
    close(st["rho"][0], exp, 1e-6, "clamped predicted density")
    close(st["psi"][1], psi, 1e-6, "lone boundary volume")


def _coupling_pair(kind, frac):
    p = base_params(mode=upf.UPF_BOUNDARY_AKINCI)
    p.surface_tension_gamma = 0.05
    p.surface_adhesion_beta = 0.01
    h = float(p.support_radius)
    x = line((0.3, 0.3, 0.3), (2, 3, 6), (frac * h,))
    typ = [2, 2] if kind == "tension" else [2, 6]
    o, st = run(p, x, typ, upf.UPF_PCISPH_RIGID)
    return p, h, f64(x), st


def pin_surface_tension():
    """Cohesion between two liquid particles at 0.7117 h (attracting) and at
    0.2617 h (repelling, where the spline is negative):
    F_i = -(x_i - x_j) gamma m^2 C(d) / d."""
    for frac, attract in ((0.7117, True), (0.2617, False)):
        p, h, x, st = _coupling_pair("tension", frac)
        m, gam = float(p.initial_mass), float(p.surface_tension_gamma)
        d = np.linalg.norm(x[0] - x[1])
        F0 = -(x[0] - x[1]) * gam * m * m * table("tension", d, h) / d
        close(st["force"][0, :3], F0, 1e-5, "tension force at %.4f h" % frac)
        close(st["force"][1, :3], -F0, 1e-5, "tension reaction at %.4f h" % frac)
        toward = np.dot(st["force"][0, :3], x[1] - x[0]) > 0
        assert toward == attract


def pin_adhesion():
    """Adhesion of a liquid particle to a lone boundary particle at 0.7117 h
    (Akinci volume Psi = 0.7 / W(0)): F = -(x_i - x_b) beta m rho0 Psi A(d) / d,
    pulling the liquid particle toward the boundary."""
    p, h, x, st = _coupling_pair("adhesion", 0.7117)
    m, beta, rho0 = float(p.initial_mass), float(p.surface_adhesion_beta), float(p.rest_density)
    psi = 0.7 / kernel("w", 0.0, h)
    d = np.linalg.norm(x[0] - x[1])
    F0 = -(x[0] - x[1]) * beta * m * rho0 * psi * table("adhesion", d, h) / d
    close(st["force"][0, :3], F0, 1e-5, "adhesion force")
    assert np.dot(st["force"][0, :3], x[1] - x[0]) > 0


def pin_spring_and_damping():
    """Two one-particle rigid bodies overlapping (distance 0.6 spacing < 2 r):
    spring -k (2 r - d) (x_j - x_i) / d pushes them apart, damping
    c (v_j - v_i) opposes the relative velocity; each body's momentum grows
    by F dt and its position by the new velocity times dt."""
    p = base_params(mode=upf.UPF_BOUNDARY_AKINCI)
    r = float(p.particle_radius)
    x = line((0.3, 0.3, 0.3), (1, 2, 2), (0.6 * float(p.spacing),))
    v = np.float32([[0.05, 0.0, -0.02], [-0.03, 0.04, 0.0]])
    masses = (0.002, 0.005)
    bodies = []
    for k in range(2):
        b = ol.Body()
        for a in range(3):
            b.pos[a] = x[k, a]
            b.vel[a] = v[k, a]
            b.linear_momentum[a] = np.float32(masses[k]) * v[k, a]
        b.mass = masses[k]
        b.quaternion[0] = 1.0
        for a in range(3):
            b.rotation[4 * a] = 1.0
            b.inv_inertia_local[4 * a] = 1.0
            b.inv_inertia_world[4 * a] = 1.0
        bodies.append(b)
    bid = np.int32([0, 1])
    rel = np.zeros((2, 4), np.float32)
    pmass = np.float32(masses)
    o, st = run(p, x, [0, 0], upf.UPF_PCISPH_RIGID, vel=v, bodies=(bodies, bid, rel, pmass))
    xi, vi = f64(x), f64(v)
    d = np.linalg.norm(xi[1] - xi[0])
    k, c, dt = float(p.spring_coefficient), float(p.damping_coefficient), float(p.dt)
    F0 = -k * (2 * r - d) * (xi[1] - xi[0]) / d + c * (vi[1] - vi[0])
    F1 = -k * (2 * r - d) * (xi[0] - xi[1]) / d + c * (vi[0] - vi[1])
    bs = o.bodies(2)
    for b, F, mk, x0, v0 in ((bs[0], F0, masses[0], xi[0], vi[0]), (bs[1], F1, masses[1], xi[1], vi[1])):
        P1 = mk * v0 + F * dt
        close(np.float64(list(b.linear_momentum)), P1, 1e-5, "body momentum")
        close(np.float64(list(b.vel)), P1 / mk, 1e-5, "body velocity")
        close(np.float64(list(b.pos)) - x0, P1 / mk * dt, 1e-4, "body displacement", atol=1e-7)
    assert np.dot(F0, xi[1] - xi[0]) < 0  # the spring dominates: apart


def pin_tait_equation_of_state():
    """WCSPH pressures of a compressed block against the Tait closed form
    p = max(0, rho0 k / 7 ((rho / rho0)^7 - 1)) on the densities of the pass."""
    p = upf.make_params(0.01, 0.000253)
    pos, typ, nf = upf.dam_break(p, 6, 6, 6, (0.12, 0.06, 0.06), lattice_rel=0.97)
    p.gas_const_wcsph = 1000.0
    o = ol.OracleSim(p, pos, None, typ, omp=False)
    o.step(1, upf.UPF_WCSPH_STATIC)
    st = o.state(pcisph=False)
    f = typ == upf.UPF_LIQUID
    rho = st["rho"][f].astype(np.float64)
    rho0, k = float(p.rest_density), float(p.gas_const_wcsph)
    exp = np.maximum(0.0, rho0 * k / 7.0 * ((rho / rho0) ** 7 - 1.0))
    assert np.count_nonzero(exp) > 10
    close(st["p"][f], exp, 2e-5, "Tait pressure")


def pin_wcsph_pair_forces():
    """WCSPH forces on three liquid particles (gaps 0.4517 h and 0.7183 h,
    rest density 300, stiffness 1000, moving): Monaghan's symmetric pressure
    force -m^2 (p_i / rho_i^2 + p_j / rho_j^2) grad W, the viscosity term
    -(v_i - v_j) (m / rho_i) (m / rho_j) mu lap W, gravity -m g; then the
    symplectic Euler update."""
    p = base_params(gravity=9.81, rho0=300.0)
    p.gas_const_wcsph = 1000.0
    h = float(p.support_radius)
    x = line((0.3, 0.3, 0.3), (6, 2, 3), (0.4517 * h, 0.7183 * h))
    v = np.float32([[0.1, 0.0, 0.0], [0.0, -0.1, 0.05], [0.0, 0.0, 0.2]])
    o, st = run(p, x, [2, 2, 2], upf.UPF_WCSPH_STATIC, vel=v)
    xi, vi = f64(x), f64(v)
    m, rho0, k, mu, g, dt = (float(p.initial_mass), float(p.rest_density), float(p.gas_const_wcsph),
                             float(p.visc_const), float(p.gravity), float(p.dt))
    n = 3
    nb = [[j for j in range(n) if j != i and np.linalg.norm(xi[i] - xi[j]) < h] for i in range(n)]
    rho = np.array([m * kernel("w", 0.0, h) + sum(m * table("w", np.linalg.norm(xi[i] - xi[j]), h) for j in nb[i])
                    for i in range(n)])
    pr = np.maximum(0.0, rho0 * k / 7.0 * ((rho / rho0) ** 7 - 1.0))
    F = np.zeros((n, 3))
    for i in range(n):
        for j in nb[i]:
            d = np.linalg.norm(xi[i] - xi[j])
            F[i] -= m * m * (pr[i] / rho[i] ** 2 + pr[j] / rho[j] ** 2) * table("grad", d, h) * (xi[i] - xi[j])
            F[i] -= (vi[i] - vi[j]) * (m / rho[i]) * (m / rho[j]) * mu * table("visc", d, h)
        F[i, 1] -= m * g
    v1 = vi + F / m * dt
    close(st["rho"][:], rho, 1e-6, "WCSPH density")
    close(st["p"][:], pr, 2e-5, "WCSPH pressure")
    close(st["force"][:, :3], F, 2e-5, "WCSPH force")
    close(st["vel"][:, :3], v1, 2e-5, "WCSPH velocity")
    close(st["pos"][:, :3] - xi, v1 * dt, 1e-4, "WCSPH displacement", atol=1e-7)
    assert len(set(np.round(pr, 3))) == 3


def pin_hydrostatic_column():
    """A 6 x 12 x 6 fluid column (spacing 0.01 m) in a box whose floor
    matches its footprint, settling under gravity with PCISPH for 0.4 s: the
    predicted density stays within 1 % of rho0 (the loop's threshold), the
    column comes to rest, the correction pressure grows with depth at the
    hydrostatic rate rho0 g (within the SPH discretisation's 30 %: the
    measured slope is 1.28 rho0 g), and in the interior layers the pressure
    force carries the weight (mean total force a small fraction of m g; the
    bottom layer rests on the floor clamp, the top layer is the free surface)."""
    p = upf.make_params(0.01, 0.001)
    pos, typ, nf = upf.dam_break(p, 6, 12, 6, (0.06, 0.24, 0.06))
    upf.density_error_factor(p)
    o = ol.OracleSim(p, pos, None, typ, omp=True)
    o.step(400, upf.UPF_PCISPH_STATIC)
    st = o.state()
    f = typ == upf.UPF_LIQUID
    K, rmax = o.stats()
    m, g, rho0 = float(p.initial_mass), float(p.gravity), float(p.rest_density)
    speed = np.linalg.norm(st["vel"][f, :3], axis=1)
    assert speed.max() < 0.1, speed.max()  # free fall over 0.4 s would reach 3.9 m/s
    assert np.all(K[-50:] < 50)
    assert (rmax[-50:].max() - rho0) / rho0 < 0.01, rmax[-50:].max()
    assert abs(st["rho"][f].astype(np.float64).mean() - rho0) / rho0 < 0.005
    ys = st["pos"][f, 1].astype(np.float64)
    layer = np.round((ys - float(p.collision_min[1])) / 0.01).astype(int)
    assert layer.min() == 0 and layer.max() <= 12
    inner = (layer >= 1) & (layer <= 10)
    slope = np.polyfit(ys[inner], st["p"][f][inner].astype(np.float64), 1)[0]
    assert 0.8 < -slope / (rho0 * g) < 1.6, -slope / (rho0 * g)
    Fy = st["force"][f, 1].astype(np.float64)
    assert abs(Fy[(layer >= 2) & (layer <= 10)].mean()) < 0.25 * m * g


PINS = [pin_pcisph_pressure_three_particles, pin_pcisph_viscosity_and_gravity, pin_free_fall_and_floor_hit,
        pin_predicted_position_clamp, pin_surface_tension, pin_adhesion, pin_spring_and_damping,
        pin_tait_equation_of_state, pin_wcsph_pair_forces, pin_hydrostatic_column]


@pytest.mark.parametrize("pin", PINS, ids=[f.__name__[4:] for f in PINS])
def test_pin(pin):
    pin()


# ------------------------------------------------------ planted mistakes
# planted mistake -> pins that must fail on it (-DORC_PLANT=k in the oracle)
PLANTS = {
    1: ("corrective pressure force with a flipped sign", ["pcisph_pressure_three_particles"]),
    2: ("corrective pressure force with m instead of m^2", ["pcisph_pressure_three_particles"]),
    3: ("corrective pressure force with p_i in place of p_j", ["pcisph_pressure_three_particles"]),
    4: ("explicit instead of symplectic Euler", ["pcisph_pressure_three_particles", "free_fall_and_floor_hit"]),
    5: ("viscosity with a flipped sign", ["pcisph_viscosity_and_gravity"]),
    6: ("floor damping +0.5 instead of -0.5", ["free_fall_and_floor_hit"]),
    7: ("no collision clamp of the predicted position", ["predicted_position_clamp"]),
    8: ("Tait exponent 6 instead of 7", ["tait_equation_of_state", "wcsph_pair_forces"]),
    9: ("surface tension with a flipped sign", ["surface_tension"]),
    10: ("adhesion with a flipped sign", ["adhesion"]),
    11: ("spring force with a flipped sign", ["spring_and_damping"]),
    12: ("damping force with a flipped sign", ["spring_and_damping"]),
    13: ("gravity force g instead of m g", ["pcisph_viscosity_and_gravity", "free_fall_and_floor_hit"]),
    14: ("predicted density without the self term", ["pcisph_pressure_three_particles"]),
    15: ("correction pressure not accumulated over iterations", ["pcisph_pressure_three_particles"]),
    16: ("WCSPH pressure force with a flipped sign", ["wcsph_pair_forces"]),
    17: ("WCSPH pressure force with p_i / rho_i^2 in place of p_j / rho_j^2", ["wcsph_pair_forces"]),
}


def _build_planted(tmpdir):
    src = os.path.join(ol.ORACLE_DIR, "upf_oracle.cpp")
    inc = os.path.join(os.path.dirname(ol.ORACLE_DIR), "include")
    procs = {}
    for k in PLANTS:
        out = os.path.join(tmpdir, "liborc_plant%d.so" % k)
        cmd = ["g++", "-std=c++17", "-O1", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
               "-Wno-unknown-pragmas", "-I" + inc, "-DORC_PLANT=%d" % k, "-o", out, src]
        procs[k] = (subprocess.Popen(cmd), out)
        if len(procs) % 8 == 0:
            for q, _ in procs.values():
                q.wait()
    libs = {}
    for k, (q, out) in procs.items():
        assert q.wait() == 0, "planted build %d failed" % k
        libs[k] = out
    return libs


def test_planted_mistakes_fail_a_pin(tmp_path):
    libs = _build_planted(str(tmp_path))
    names = {f.__name__[4:]: f for f in PINS}
    report = []
    for k, (what, must_fail) in PLANTS.items():
        failed = []
        with ol.planted(libs[k]):
            for name in must_fail:
                try:
                    names[name]()
                except (AssertionError, RuntimeError):
                    failed.append(name)
        report.append((k, what, failed))
        assert failed == must_fail, "plant %d (%s): pins %s passed" % (k, what, sorted(set(must_fail) - set(failed)))
    # the unplanted oracle passes all of them
    for f in PINS:
        f()
