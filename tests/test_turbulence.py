"""Decaying-turbulence initial conditions (CPU): the jittered lattice, the
solenoidal Gaussian velocity field and its periodic sampling.

The velocity field is pinned by what defines it: zero divergence
(k . v(k) = 0 for every mode), no power outside 2 <= |k| <= 8, a shell
power spectrum |v(k)|^2 ~ k^-4 (fitted slope), Gaussian one-point
statistics (kurtosis 3) and unit rms."""
import numpy as np
import pytest

from meshlesshydro_amd import turbulence as T


def _modes(ng):
    m = np.fft.fftfreq(ng) * ng
    mz = np.fft.rfftfreq(ng) * ng
    return np.meshgrid(m, m, mz, indexing="ij")


def test_solenoidal_field_properties():
    ng = 48
    v = T.solenoidal_field(ng, kmin=2, kmax=8, slope=-4.0, seed=7)
    assert v.shape == (3, ng, ng, ng)
    assert abs(np.sqrt(np.mean(np.sum(v * v, axis=0))) - 1.0) < 1e-12
    V = np.fft.rfftn(v, axes=(1, 2, 3))
    KX, KY, KZ = _modes(ng)
    K = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    div = KX * V[0] + KY * V[1] + KZ * V[2]
    amp = np.sqrt(np.abs(V[0]) ** 2 + np.abs(V[1]) ** 2 + np.abs(V[2]) ** 2)
    assert np.max(np.abs(div)) < 1e-10 * np.max(K * amp)
    power = amp ** 2
    band = (K >= 2) & (K <= 8)
    assert power[~band].sum() < 1e-20 * power[band].sum()
    # spectrum: log |v(k)|^2 against log |k| over the band's modes (the
    # per-mode scatter of a Gaussian field only shifts the intercept)
    slope = np.polyfit(np.log(K[band]), np.log(power[band]), 1)[0]
    assert abs(slope + 4.0) < 0.15, slope
    # Gaussian components
    for c in range(3):
        x = v[c].ravel()
        kurt = np.mean((x - x.mean()) ** 4) / np.var(x) ** 2
        assert abs(kurt - 3.0) < 0.15, kurt


def test_sample_periodic_reproduces_grid_and_wraps():
    ng = 16
    rng = np.random.default_rng(1)
    f = rng.standard_normal((2, ng, ng, ng))
    i, j, k = rng.integers(0, ng, (3, 50))
    out = T.sample_periodic(f, i / ng, j / ng, k / ng)
    assert np.allclose(out[0], f[0, i, j, k], atol=1e-14, rtol=0)
    # a field linear inside one cell is interpolated exactly; across the
    # periodic seam the interpolation uses points ng-1 and 0
    g = np.zeros((1, ng, ng, ng))
    g[0, ng - 1] = 1.0
    x = np.array([(ng - 1 + 0.25) / ng])
    assert np.allclose(T.sample_periodic(g, x, np.zeros(1), np.zeros(1)), 0.75)
    c = np.full((1, ng, ng, ng), 2.5)
    xs = rng.uniform(0, 1, (3, 100))
    assert np.allclose(T.sample_periodic(c, *xs), 2.5)


def test_jittered_lattice_and_kernel_size():
    n = 12
    x, y, z = T.jittered_lattice(n, seed=4242)
    x2, _, _ = T.jittered_lattice(n, seed=4242)
    assert np.array_equal(x, x2)
    assert len(x) == n ** 3 and np.all((x >= 0) & (x < 1)) and np.all((z >= 0) & (z < 1))
    c = (np.arange(n) + 0.5) / n
    base = np.tile(c, n * n)
    d = (z - base + 0.5) % 1.0 - 0.5
    assert np.max(np.abs(d)) <= 0.3 / n + 1e-15
    h = T.kernel_size(n)
    assert abs(4.0 / 3.0 * np.pi * (h * n) ** 3 - 64.0) < 1e-9
