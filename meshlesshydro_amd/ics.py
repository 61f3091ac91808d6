"""Initial conditions for the benchmark and parity workloads.

* Kelvin-Helmholtz shear layer: the formulas of
  testcases/kelvin-helmholtz/generateIC.py:10-119 (regular-grid branch
  :75-86, x outer / y inner; random branch :88-93 with seed 6102003),
  restated here with the same numpy operations.
* 3D point-energy blast on a cell-centred lattice in the periodic unit cube:
  rho = 1, P = 1e-6 background, E = 1 deposited on the particles within
  h_dep = 2 dx of the centre with cubic-spline weights (the 32 particles at
  r = sqrt(.75) dx and sqrt(2.75) dx).
* Random Riemann face states (microbenchmark): rho_L, rho_R log-uniform
  [1e-3, 1e3], P log-uniform [1e-6, 1e3], u_L = 0, u_R = U[-10,10] c_s with
  c_s = (c_L + c_R)/2.
"""
import numpy as np

GAMMA = 5.0 / 3.0


# --- Kelvin-Helmholtz (generateIC.py:10-59) --------------------------------
_rho2, _rho1, _Dy, _P = 2.0, 1.0, 0.025, 5.0 / 2.0
_Drho = (_rho1 - _rho2) / 2.0
_dvy0, _v1, _v2 = 0.01, 0.5, -0.5
_Dv = (_v1 - _v2) / 2.0


def _kh_vx(y):
    v = np.empty(len(y))
    m1 = y < .25
    v[m1] = _v1 - _Dv * np.exp((y[m1] - .25) / _Dy)
    m2 = (.25 <= y) & (y < .5)
    v[m2] = _v2 + _Dv * np.exp((.25 - y[m2]) / _Dy)
    m3 = (.5 <= y) & (y <= .75)
    v[m3] = _v2 + _Dv * np.exp((y[m3] - .75) / _Dy)
    m4 = .75 < y
    v[m4] = _v1 - _Dv * np.exp((.75 - y[m4]) / _Dy)
    return v


def _kh_rho(y):
    d = np.empty(len(y))
    m1 = y < .25
    d[m1] = _rho1 - _Drho * np.exp((y[m1] - .25) / _Dy)
    m2 = (.25 <= y) & (y < .5)
    d[m2] = _rho2 + _Drho * np.exp((.25 - y[m2]) / _Dy)
    m3 = (.5 <= y) & (y <= .75)
    d[m3] = _rho2 + _Drho * np.exp((y[m3] - .75) / _Dy)
    m4 = .75 < y
    d[m4] = _rho1 - _Drho * np.exp((.75 - y[m4]) / _Dy)
    return d


def kelvin_helmholtz(n_side, regular=True, gamma=GAMMA, seed=6102003):
    """2D shear layer with n_side**2 particles in [0,1)^2.

    Returns dict x, y, vx, vy, m, u (float64 arrays) in particle order."""
    N = n_side * n_side
    if regular:
        xv = np.linspace(0., 1., n_side, endpoint=False)
        pos = np.empty((N, 2))
        pos[:, 0] = np.repeat(xv, n_side)
        pos[:, 1] = np.tile(xv, n_side)
    else:
        rng = np.random.default_rng(seed)
        pos = rng.random(size=(N, 2))
    vel = np.empty(pos.shape)
    vel[:, 0] = _kh_vx(pos[:, 1])
    vel[:, 1] = _dvy0 * np.sin(4. * np.pi * pos[:, 0])
    rho = _kh_rho(pos[:, 1])
    m = rho / N
    u = _P / ((gamma - 1.) * rho)
    return dict(x=pos[:, 0].copy(), y=pos[:, 1].copy(), vx=vel[:, 0].copy(),
                vy=vel[:, 1].copy(), m=m, u=u)


def cubic_spline(dim, r, h):
    """Kernel::cubicSpline (Particles.cpp:7-24), vectorised."""
    h2 = h / 2.
    sigma = 10. / (7. * np.pi * h2 * h2) if dim == 2 else 1. / (np.pi * h2 * h2 * h2)
    q = np.asarray(r, dtype=np.float64) / h2
    w = np.zeros_like(q)
    a = (0. <= q) & (q <= 1.)
    w[a] = sigma * (1. - 3. / 2. * q[a] * q[a] * (1. - q[a] / 2.))
    b = (1. < q) & (q < 2.)
    w[b] = sigma / 4. * (2. - q[b]) ** 3
    return w


def blast_3d(n_side, rho0=1.0, P0=1e-6, E=1.0, gamma=GAMMA, h_dep_dx=2.0):
    """3D point-energy blast on an n_side^3 cell-centred lattice."""
    N = n_side ** 3
    dx = 1.0 / n_side
    c = (np.arange(n_side) + 0.5) * dx
    X = np.repeat(c, n_side * n_side)
    Y = np.tile(np.repeat(c, n_side), n_side)
    Z = np.tile(c, n_side * n_side)
    m = np.full(N, rho0 / N)
    u = np.full(N, P0 / ((gamma - 1.) * rho0))
    r = np.sqrt((X - .5) ** 2 + (Y - .5) ** 2 + (Z - .5) ** 2)
    w = cubic_spline(3, r, h_dep_dx * dx)
    sel = w > 0
    u[sel] += E * (w[sel] / w[sel].sum()) / m[sel]
    zero = np.zeros(N)
    return dict(x=X, y=Y, z=Z, vx=zero.copy(), vy=zero.copy(), vz=zero.copy(),
                m=m, u=u)


def blast_kernel_size(n_side):
    """h = 2.4 dx: 56 lattice neighbours in 3D."""
    return 2.4 / n_side


def riemann_states(n, seed=20261202, gamma=GAMMA):
    """Random face states for the Riemann microbenchmark."""
    rng = np.random.default_rng(seed)
    rl = 10.0 ** rng.uniform(-3, 3, n)
    rr = 10.0 ** rng.uniform(-3, 3, n)
    pl = 10.0 ** rng.uniform(-6, 3, n)
    pr = 10.0 ** rng.uniform(-6, 3, n)
    cs = 0.5 * (np.sqrt(gamma * pl / rl) + np.sqrt(gamma * pr / rr))
    ul = np.zeros(n)
    ur = rng.uniform(-10, 10, n) * cs
    return rl, ul, pl, rr, ur, pr
