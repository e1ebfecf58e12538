"""Pins of the oracle's FP64 conservation sums (oracle.conservation).

  kinetic / potential energy and momentum   closed forms for a few particles
  inactive particles                        excluded from every sum
  EOS internal energy                       zero at rest density; equals the
        numerical integral of p / rho^2 from rho0 (scipy quad); k eps^2 / 2 to
        leading order for rho = rho0 (1 + eps)
"""
import numpy as np
from scipy.integrate import quad


def test_closed_forms_and_inactive(orc):
    pos = np.array([[0.1, 0.2, 0.3, 1.0], [0.5, 0.5, 0.5, -1.0], [0.2, 0.4, 0.1, 1.0]], np.float32)
    vel = np.array([[1.0, 2.0, 3.0, 0.0], [9.0, 9.0, 9.0, 0.0], [-1.0, 0.0, 0.5, 0.0]], np.float32)
    rho_p = np.array([[1000.0, 0.0], [1.0, 0.0], [1000.0, 0.0]], np.float32)
    m, g, s = 0.002, (0.0, -9.81, 0.0), 2.0
    out = orc.conservation(pos, vel, pos, rho_p, m, g, s, 1000.0, 3.0)
    v = vel[[0, 2], :3].astype(np.float64)
    assert out[0] == 2 and out[1] == 2 * m
    assert np.allclose(out[2:5], m * v.sum(0), rtol=1e-15)
    assert np.isclose(out[5], 0.5 * m * (14.0 + 1.25), rtol=1e-15)
    assert np.isclose(out[6], m * 9.81 * s * (np.float32(0.2) + np.float32(0.4)), rtol=1e-12)
    assert out[7] == 0.0  # both active particles sit at rest density


def test_eos_internal_energy(orc):
    k, r0, m = 3.0, 1000.0, 0.125
    for rho in (900.0, 1001.0, 1100.0, 2.0e3):
        pos = np.array([[0.0, 0.0, 0.0, 1.0]], np.float32)
        out = orc.conservation(pos, np.zeros((1, 4), np.float32), pos, np.array([[rho, 0.0]], np.float32), m,
                               (0, 0, 0), 1.0, r0, k)
        ref = m * quad(lambda r: k * (r - r0) / (r * r), r0, float(np.float32(rho)), epsabs=0, epsrel=1e-13)[0]
        assert np.isclose(out[7], ref, rtol=1e-9, atol=1e-15)
    eps = 1e-3
    pos = np.array([[0.0, 0.0, 0.0, 1.0]], np.float32)
    rho = np.float32(r0 * (1 + eps))
    e = orc.conservation(pos, np.zeros((1, 4), np.float32), pos, np.array([[rho, 0.0]], np.float32), 1.0,
                         (0, 0, 0), 1.0, r0, k)[7]
    e_lead = k * ((float(rho) - r0) / r0) ** 2 / 2
    assert abs(e - e_lead) < 2 * k * eps ** 3
