"""Original SPH (physicsType 'o': CpuParticlePhysicsStaticBoundariesSPH with
CalculatePressureSPH, src/cpu/UnifiedPhysics.cpp:4180-4202,2479-2489) in the
oracle: the density pass is WCSPH's, the pressure the linear equation of
state p = max(0, k (rho - rho0)) with k = fluidGasConst."""
import numpy as np

import oracle_lib as ol
import scenes
import unifiedparticleframeworkcuda_amd as upf


def test_sph_pressure_is_linear_in_the_density_of_the_wcsph_pass():
    p, pos, typ, nf = scenes.dam_break(8)
    p.dt = 0.000253
    f = typ == upf.UPF_LIQUID
    a = ol.OracleSim(p, pos, None, typ, omp=False)
    b = ol.OracleSim(p, pos, None, typ, omp=False)
    a.step(1, upf.UPF_SPH_STATIC)
    b.step(1, upf.UPF_WCSPH_STATIC)
    sa, sb = a.state(pcisph=False), b.state(pcisph=False)
    assert np.array_equal(sa["rho"][f], sb["rho"][f])  # same density pass
    rho0 = np.float32(p.rest_density)
    k = np.float32(p.gas_const_sph)
    assert k == np.float32(1000.0)
    rho = sa["rho"][f]
    exp = np.maximum(np.float32(0), np.float32(rho0 * k) * ((rho / rho0).astype(np.float32) - np.float32(1)))
    assert np.array_equal(sa["p"][f], exp.astype(np.float32))
    assert np.any(sa["p"][f] > 0) and np.any(sa["p"][f] == 0)
    # k (rho - rho0) up to the rounding of rho / rho0 (one ulp of 1, times rho0 k)
    ulp = float(rho0) * float(k) * 2.0 ** -23
    assert np.allclose(sa["p"][f], np.maximum(0.0, 1000.0 * (rho.astype(np.float64) - 1000.0)), rtol=0, atol=2 * ulp)
    assert not np.array_equal(sa["pos"][f], sb["pos"][f])  # the equations of state differ
