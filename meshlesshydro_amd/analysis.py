"""Analysis of runs: conservation errors from snapshots and the Sedov-Taylor
point-blast solution for the 3D blast.

Host-side numpy/scipy, like the reference's own analysis scripts; nothing
here is on the time-step path.

* `conservation_series` follows testcases/kelvin-helmholtz/
  conservationPlotter.py:33-45: the absolute deviation of total mass,
  energy and momentum from the first snapshot's values.
* `SedovTaylor` is the self-similar strong point explosion in a uniform
  medium (Sedov 1959; Taylor 1950; Landau & Lifshitz, Fluid Mechanics
  par. 106), the solution testcases/sedov/analytical.py evaluates.  It is
  obtained here by integrating the similarity form of the Euler equations
  inward from the strong-shock jump conditions, not from a parametric
  closed form.
"""
import glob
import math
import os

import numpy as np
from scipy.integrate import simpson, solve_ivp

from . import host

__all__ = ["conservation_series", "SedovTaylor", "radial_profile", "sedov_compare"]


def conservation_series(out_dir):
    """Absolute conservation errors over the snapshots `out_dir/*.h5`.

    Returns a dict of numpy arrays sorted by time: `time`, `mass`,
    `energy`, `momentum_x`, `momentum_y` and, for 3D snapshots,
    `momentum_z`; each error is |Q(t) - Q(t_0)| with t_0 the first
    snapshot in file-name order (the writer numbers them %06d).  The
    reference script takes its reference values from whichever file its
    directory listing yields first; file-name order makes that the initial
    snapshot.
    """
    files = sorted(glob.glob(os.path.join(out_dir, "*.h5")))
    if not files:
        raise FileNotFoundError(f"no .h5 snapshots in {out_dir}")
    keys = [("mass", "totalMass"), ("energy", "energy"), ("momentum_x", "xMomentum"),
            ("momentum_y", "yMomentum")]
    try:
        host.read_dataset(files[0], "zMomentum")
        keys.append(("momentum_z", "zMomentum"))
    except host.HostError:
        pass
    rows = []
    for f in files:
        rows.append([float(host.read_dataset(f, "time")[0])] +
                    [float(host.read_dataset(f, ds)[0]) for _, ds in keys])
    a = np.array(rows)
    out = {"time": a[:, 0]}
    for c, (name, _) in enumerate(keys, start=1):
        out[name] = np.abs(a[:, c] - a[0, c])
    order = np.argsort(out["time"], kind="stable")
    return {k: v[order] for k, v in out.items()}


class SedovTaylor:
    """Strong point explosion of energy E in a uniform gas of density rho0
    (pressure neglected) in nu = 1, 2 or 3 dimensions (planar, cylindrical,
    spherical).

    Similarity form: R(t) = xi0 (E t^2 / rho0)^(1/(nu+2)), Rdot = 2R/((nu+2)t),
    eta = r/R and

        v = Rdot f(eta),   rho = rho0 g(eta),   p = rho0 Rdot^2 h(eta).

    Substituting into the Euler equations gives, with u = f - eta,
    (integrated in f, ln g and ln h)

        u g' + g f'               = -(nu-1) g f / eta       (mass)
        u f' + h'/g               = (nu/2) f                (momentum)
        u (h'/h - gamma g'/g)     = nu                      (entropy)

    integrated from the strong-shock state at eta = 1,
    f = h = 2/(gamma+1), g = (gamma+1)/(gamma-1), towards the centre.
    Inward integration is unstable near the centre (the homogeneous
    solution f ~ eta^-(nu-1) grows), so it stops at eta_min = 0.01 and the
    centre follows the asymptotic solution matched there: f ~ eta (f/eta ->
    1/gamma), g ~ eta^(nu/(gamma-1)), h constant.  xi0
    follows from the energy integral
    E = rho0 Rdot^2 R^nu S_nu int_0^1 (g f^2/2 + h/(gamma-1)) eta^(nu-1) deta,
    with S_nu = 2, 2 pi, 4 pi.
    """

    def __init__(self, E=1.0, rho0=1.0, gamma=5.0 / 3.0, nu=3, eta_min=1e-2, n_eta=4001):
        if nu not in (1, 2, 3):
            raise ValueError("nu must be 1, 2 or 3")
        self.E, self.rho0, self.gamma, self.nu = float(E), float(rho0), float(gamma), int(nu)
        g_ = self.gamma
        nu_ = self.nu

        def rhs(eta, y):
            # unknowns f, ln g, ln h: g falls like a high power of eta near
            # the centre, so its logarithm keeps the system well scaled
            f, lg, lh = y
            u = f - eta
            hg = math.exp(lh - lg)
            # (f', (ln g)', (ln h)') from the three equations divided by g, h
            A = np.array([[1.0, u, 0.0],
                          [u, 0.0, hg],
                          [0.0, -g_ * u, u]])
            b = np.array([-(nu_ - 1) * f / eta, 0.5 * nu_ * f, float(nu_)])
            return np.linalg.solve(A, b)

        y1 = [2.0 / (g_ + 1.0), math.log((g_ + 1.0) / (g_ - 1.0)), math.log(2.0 / (g_ + 1.0))]
        eta = np.linspace(1.0, eta_min, n_eta)
        sol = solve_ivp(rhs, (1.0, eta_min), y1, t_eval=eta, method="LSODA",
                        rtol=1e-11, atol=1e-13)
        if not sol.success:
            raise RuntimeError(f"Sedov similarity integration failed: {sol.message}")
        # ascending eta for interpolation
        self.eta = sol.t[::-1].copy()
        self.f = sol.y[0][::-1].copy()
        self.g = np.exp(sol.y[1][::-1])
        self.h = np.exp(sol.y[2][::-1])
        s_nu = {1: 2.0, 2: 2.0 * math.pi, 3: 4.0 * math.pi}[nu_]
        self.s_nu = s_nu
        self.g_exp = nu_ / (g_ - 1.0)
        integrand = (0.5 * self.g * self.f ** 2 + self.h / (g_ - 1.0)) * self.eta ** (nu_ - 1)
        e0 = self.eta[0]
        centre = (self.h[0] / (g_ - 1.0) * e0 ** nu_ / nu_ +
                  0.5 * self.g[0] * self.f[0] ** 2 * e0 ** nu_ / (self.g_exp + nu_ + 2.0))
        self.energy_integral = float(simpson(integrand, x=self.eta)) + centre
        self.delta = 2.0 / (nu_ + 2.0)
        self.xi0 = (self.delta ** 2 * s_nu * self.energy_integral) ** (-1.0 / (nu_ + 2.0))

    def mass_integral(self):
        """int_0^1 g eta^(nu-1) deta: equals 1/nu when the swept-up mass is
        conserved (a check the integration does not impose)."""
        e0 = self.eta[0]
        centre = self.g[0] * e0 ** self.nu / (self.g_exp + self.nu)
        return float(simpson(self.g * self.eta ** (self.nu - 1), x=self.eta)) + centre

    def shock_radius(self, t):
        return self.xi0 * (self.E * t * t / self.rho0) ** (1.0 / (self.nu + 2.0))

    def shock_speed(self, t):
        return self.delta * self.shock_radius(t) / t

    def time_for_radius(self, R):
        return math.sqrt((R / self.xi0) ** (self.nu + 2.0) * self.rho0 / self.E)

    def _interp(self, arr, eta, outside, centre_power):
        eta = np.asarray(eta, dtype=float)
        out = np.interp(eta, self.eta, arr)
        # below eta_min: the centre asymptotics, arr ~ eta^centre_power
        e0 = self.eta[0]
        lo = eta < e0
        out = np.where(lo, arr[0] * (np.maximum(eta, 0.0) / e0) ** centre_power, out)
        return np.where(eta > 1.0, outside, out)

    def density(self, r, t):
        eta = np.asarray(r, dtype=float) / self.shock_radius(t)
        return self.rho0 * self._interp(self.g, eta, 1.0, self.g_exp)

    def velocity(self, r, t):
        eta = np.asarray(r, dtype=float) / self.shock_radius(t)
        return self.shock_speed(t) * self._interp(self.f, eta, 0.0, 1.0)

    def pressure(self, r, t):
        eta = np.asarray(r, dtype=float) / self.shock_radius(t)
        return self.rho0 * self.shock_speed(t) ** 2 * self._interp(self.h, eta, 0.0, 0.0)


def radial_profile(x, values, center, r_max, nbins, weights=None):
    """Mean of `values` in `nbins` radial shells of [0, r_max) about
    `center` (weighted by `weights` when given).  x is [N][D] or a tuple of
    coordinate arrays.  Returns (shell centres, means, counts)."""
    if isinstance(x, (tuple, list)):
        x = np.stack([np.asarray(c) for c in x], axis=1)
    r = np.sqrt(((np.asarray(x) - np.asarray(center)) ** 2).sum(axis=1))
    edges = np.linspace(0.0, r_max, nbins + 1)
    idx = np.digitize(r, edges) - 1
    ok = (idx >= 0) & (idx < nbins)
    w = np.ones_like(r) if weights is None else np.asarray(weights, dtype=float)
    sw = np.bincount(idx[ok], weights=w[ok], minlength=nbins)
    sv = np.bincount(idx[ok], weights=(w * np.asarray(values))[ok], minlength=nbins)
    cnt = np.bincount(idx[ok], minlength=nbins)
    mean = np.divide(sv, sw, out=np.full(nbins, np.nan), where=sw > 0)
    return 0.5 * (edges[1:] + edges[:-1]), mean, cnt


def sedov_compare(x, rho, P, t, E=1.0, rho0=1.0, gamma=5.0 / 3.0, center=(0.5, 0.5, 0.5),
                  r_max=None, nbins=64, solution=None):
    """Compare a 3D blast with the Sedov-Taylor solution at time t.

    Returns the analytic shock radius, the radius of the largest shell-mean
    density (the numerical shock), and the L1 errors of the shell-mean
    density and pressure over shells inside 1.5 R_s, relative to the
    analytic profiles' L1 norms."""
    sol = solution or SedovTaylor(E, rho0, gamma, nu=3)
    Rs = sol.shock_radius(t)
    r_max = r_max or 1.5 * Rs
    rc, rho_m, cnt = radial_profile(x, rho, center, r_max, nbins)
    _, p_m, _ = radial_profile(x, P, center, r_max, nbins)
    ok = cnt > 0
    rho_a = sol.density(rc, t)
    p_a = sol.pressure(rc, t)
    l1_rho = float(np.abs(rho_m[ok] - rho_a[ok]).sum() / np.abs(rho_a[ok]).sum())
    l1_p = float(np.abs(p_m[ok] - p_a[ok]).sum() / np.abs(p_a[ok]).sum())
    k = int(np.nanargmax(np.where(ok, rho_m, -np.inf)))
    return {"t": float(t), "shock_radius": float(Rs), "shock_radius_measured": float(rc[k]),
            "peak_density": float(rho_m[k]), "peak_density_analytic": float(sol.density(Rs * (1 - 1e-12), t)),
            "L1_rel_rho_profile": l1_rho, "L1_rel_P_profile": l1_p, "xi0": sol.xi0,
            "shells": int(ok.sum())}
