"""Analysis counterparts: the Sedov-Taylor solution and the conservation
series (CPU).

The Sedov-Taylor solution is obtained by integrating the similarity ODEs.
It is pinned by:
* published energy constants alpha (R = (E t^2 / (alpha rho0))^(1/(nu+2)));
* swept-up mass conservation, which the integration does not impose;
* the strong-shock jump conditions and the centre asymptotics
  (v/r -> Rdot/(gamma R)), which come out of the integration rather than
  being put in;
* the blast energy recovered from the dimensional profiles.
"""
import math
import os

import numpy as np
import pytest

from meshlesshydro_amd import analysis, host


@pytest.mark.parametrize("gamma,nu,alpha", [
    # Kamm & Timmes, "On efficient generation of numerically robust Sedov
    # solutions", LA-UR-07-2849 (2007): standard gamma = 1.4 cases
    (1.4, 3, 0.851072),
    (1.4, 2, 0.984074),
])
def test_sedov_published_alpha(gamma, nu, alpha):
    s = analysis.SedovTaylor(gamma=gamma, nu=nu)
    assert abs(s.xi0 ** -(nu + 2) - alpha) < 2e-6


def test_sedov_xi0_gamma_five_thirds():
    # xi0 = 1.15167 for the spherical gamma = 5/3 blast (e.g. the value used
    # for Sedov tests of SPH and moving-mesh codes)
    s = analysis.SedovTaylor(gamma=5.0 / 3.0, nu=3)
    assert abs(s.xi0 - 1.15167) < 2e-5


@pytest.mark.parametrize("gamma,nu", [(5.0 / 3.0, 3), (1.4, 3), (1.4, 2), (1.4, 1), (1.2, 3)])
def test_sedov_mass_conservation(gamma, nu):
    s = analysis.SedovTaylor(gamma=gamma, nu=nu)
    assert abs(s.mass_integral() * nu - 1.0) < 1e-8


def test_sedov_jumps_centre_and_energy():
    E, rho0, gamma, t = 2.5, 0.7, 5.0 / 3.0, 0.3
    s = analysis.SedovTaylor(E=E, rho0=rho0, gamma=gamma, nu=3)
    R, U = s.shock_radius(t), s.shock_speed(t)
    assert abs(R - s.xi0 * (E * t * t / rho0) ** 0.2) < 1e-14
    assert abs(s.time_for_radius(R) - t) < 1e-12
    below = R * (1 - 1e-12)
    assert abs(s.density(below, t) - rho0 * (gamma + 1) / (gamma - 1)) < 1e-9
    assert abs(s.velocity(below, t) - 2 * U / (gamma + 1)) < 1e-9 * U
    assert abs(s.pressure(below, t) - 2 * rho0 * U * U / (gamma + 1)) < 1e-9 * rho0 * U * U
    above = R * (1 + 1e-9)
    assert s.density(above, t) == rho0 and s.velocity(above, t) == 0 and s.pressure(above, t) == 0
    # centre: homologous expansion v = r Rdot/(gamma R) (from the integration,
    # at radii where it still runs), density ~ r^(3/(gamma-1)), flat pressure
    for eta in (0.02, 0.05):
        r = eta * R
        assert abs(s.velocity(r, t) / r - U / (gamma * R)) < 1e-5 * U / R
    r1, r2 = 0.03 * R, 0.06 * R
    assert abs(math.log(s.density(r2, t) / s.density(r1, t)) / math.log(2) - 4.5) < 1e-3
    assert abs(s.pressure(0.02 * R, t) / s.pressure(0.05 * R, t) - 1) < 1e-4
    # energy of the dimensional profiles
    r = np.linspace(0, R, 200001)
    rho, v, p = s.density(r, t), s.velocity(r, t), s.pressure(r, t)
    e = 4 * math.pi * np.trapezoid((0.5 * rho * v * v + p / (gamma - 1)) * r * r, r)
    assert abs(e / E - 1) < 1e-5


def test_radial_profile_and_compare_on_analytic_samples():
    rng = np.random.default_rng(7)
    t = 0.02
    s = analysis.SedovTaylor()
    R = s.shock_radius(t)
    x = rng.uniform(0.5 - 1.5 * R, 0.5 + 1.5 * R, size=(400000, 3))
    r = np.sqrt(((x - 0.5) ** 2).sum(1))
    rc, mean, cnt = analysis.radial_profile(x, r, (0.5, 0.5, 0.5), 1.5 * R, 30)
    ok = cnt > 0
    assert np.all(np.abs(mean[ok] - rc[ok]) < 1.5 * R / 30)
    c = analysis.sedov_compare(x, s.density(r, t), s.pressure(r, t), t, solution=s, nbins=48)
    assert abs(c["shock_radius"] - R) < 1e-15
    assert abs(c["shock_radius_measured"] - R) <= 1.5 * R / 48
    assert c["L1_rel_rho_profile"] < 0.05 and c["L1_rel_P_profile"] < 0.05


def _snap(n, dim, time, mass, energy, mom):
    z = np.zeros(n)
    return dict(time=time, totalMass=mass, energy=energy, momentum=np.array(mom[:dim]),
                rho=z + 1, m=z + mass / n, u=z + 1, P=z + 1, noi=np.zeros(n, np.int32) + 44,
                x=np.zeros((n, dim)), v=np.zeros((n, dim)), rhoGrad=np.zeros((n, dim)))


@pytest.mark.parametrize("dim", [2, 3])
def test_conservation_series(tmp_path, dim):
    vals = [(0.0, 1.0, 2.5, (0.1, -0.2, 0.3)),
            (0.1, 1.0 + 3e-15, 2.5 - 1e-12, (0.1 + 1e-17, -0.2, 0.3 - 2e-16)),
            (0.2, 1.0 - 1e-14, 2.5 + 4e-12, (0.1, -0.2 + 5e-17, 0.3))]
    # written out of time order on purpose: file names give the order
    for k in (2, 0, 1):
        t, m, e, p = vals[k]
        host.write_snapshot(os.path.join(tmp_path, "%06d.h5" % k), _snap(5, dim, t, m, e, p), dim)
    c = analysis.conservation_series(str(tmp_path))
    assert np.array_equal(c["time"], [0.0, 0.1, 0.2])
    assert np.array_equal(c["mass"], [0.0, abs(vals[1][1] - 1.0), abs(vals[2][1] - 1.0)])
    assert np.array_equal(c["energy"], [0.0, abs(vals[1][2] - 2.5), abs(vals[2][2] - 2.5)])
    for a, name in enumerate(("momentum_x", "momentum_y", "momentum_z")[:dim]):
        assert np.array_equal(c[name], [abs(vals[k][3][a] - vals[0][3][a]) for k in range(3)])
    assert ("momentum_z" in c) == (dim == 3)
