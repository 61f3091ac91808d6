"""CPU checks of the oracle's per-method stage functions, each against a
property fixed independently of the code it checks:

* gradient(f) (Particles.cpp:1257-1270, 2457-2502) of a linear field is the
  field's exact gradient (first-order consistency of the least-squares
  gradient), and gradient(rho) equals the unlimited density gradient of the
  gradient stage bitwise (same psi-tilde, same summation order);
* slopeLimiter (:1313-1444) run after compPsijTilde + gradient equals the
  fused gradient stage bitwise, and limiting a second time is refused;
* the per-pair arrays: A_ij of a pair is minus that of its reverse pair
  bitwise (IEEE subtraction is antisymmetric), F_ij = -F_ji as stored, and
  W_R / W_L of a real pair are W_L / W_R of its reverse pair bitwise with
  the midpoint quadrature point (FIRST_ORDER_QUAD_POINT 1); for an image
  pair to round-off (the image coordinates round differently per side);
* checkFluxSymmetry (:2888-2976) counts match a numpy restatement of its
  one-sided test over the downloaded arrays, with ENFORCE_FLUX_SYM and
  without (where each side solves its own pair).
"""
import numpy as np

from cases import jitter3d, kh_random


def _oracle_at_gradients(oracle_mod, p, s, limit=True):
    o = oracle_mod.Oracle(p, s)
    o.neighbours()
    o.density()
    if limit:
        o.gradients()
    else:
        o.psi_gradients()
    return o


def test_gradient_field_linear_exact(oracle_mod):
    for p, s in (kh_random(32), jitter3d(16)):
        o = _oracle_at_gradients(oracle_mod, p, s, limit=False)
        a = np.array([0.3, -1.7, 0.9])[: p.dim]
        X = np.stack([s[k] for k in ("x", "y", "z")[: p.dim]], 1)
        # a linear field without the periodic wrap: keep to particles whose
        # neighbourhood does not cross the box boundary
        f = X @ a + 2.0
        g = o.gradient_field(f)
        inner = np.all((X > 2 * p.kernel_size) & (X < 1 - 2 * p.kernel_size), axis=1)
        assert inner.sum() > 20
        assert np.max(np.abs(g[inner] - a)) < 1e-10


def test_gradient_field_equals_gradient_stage(oracle_mod):
    p, s = jitter3d(8)
    o = _oracle_at_gradients(oracle_mod, p, s, limit=False)
    rho, P, _ = o.get_density()
    G, _ = o.get_gradients()
    assert np.array_equal(o.gradient_field(rho), G[0])
    assert np.array_equal(o.gradient_field(P), G[p.dim + 1])
    assert np.array_equal(o.gradient_field(o.state()["vx"]), G[1])


def test_split_limiter_equals_fused(oracle_mod):
    for p, s in (kh_random(32), jitter3d(8)):
        fused = _oracle_at_gradients(oracle_mod, p, s, limit=True)
        split = _oracle_at_gradients(oracle_mod, p, s, limit=False)
        unl, _ = split.get_gradients()
        split.limit()
        lim, _ = split.get_gradients()
        want, _ = fused.get_gradients()
        assert np.array_equal(lim, want)
        assert not np.array_equal(unl, want)  # the limiter did act
        try:
            split.limit()
            raise AssertionError("a second limit must be refused")
        except oracle_mod.OracleError:
            pass


def _pairs(o):
    noi, nog, nbr = o.get_neighbours()
    idx = nbr & ((1 << 27) - 1)
    code = nbr >> 27
    return noi, nog, idx, code


def _reverse(noi, nog, idx, code, i, n):
    """Reverse entry of (i, n) by the reference's search (real: the entry
    i; image: the first image entry with parent i)."""
    j = idx[i, n]
    if n < noi[i]:
        m = np.nonzero(idx[j, : noi[j]] == i)[0]
    else:
        m = noi[j] + np.nonzero(idx[j, noi[j]: noi[j] + nog[j]] == i)[0]
    return j, (int(m[0]) if len(m) else -1)


def test_face_arrays_antisymmetry_and_symmetry_count(oracle_mod):
    for flux_sym in (1, 0):
        p, s = jitter3d(8, enforce_flux_sym=flux_sym)
        o = _oracle_at_gradients(oracle_mod, p, s)
        dt = o.timestep()
        o.fluxes(dt)
        fa = o.get_face_arrays(dt)
        noi, nog, idx, code = _pairs(o)
        asym = miss = 0
        for i in range(o.N):
            for n in range(noi[i] + nog[i]):
                j, m = _reverse(noi, nog, idx, code, i, n)
                if m < 0:
                    miss += 1
                    continue
                assert np.array_equal(fa["A"][i, n], -fa["A"][j, m])
                if n < noi[i]:
                    assert np.array_equal(fa["WR"][i, n], fa["WL"][j, m])
                else:  # image coordinates round differently from each side
                    assert np.allclose(fa["WR"][i, n], fa["WL"][j, m], rtol=1e-12, atol=1e-14)
                if flux_sym:
                    assert np.array_equal(fa["F"][i, n], -fa["F"][j, m])
                nc = 4 if n >= noi[i] else p.dim + 2
                asym += int(np.any(fa["F"][j, m, :nc] + fa["F"][i, n, :nc] > 1e-20))
        got = o.check_flux_symmetry(1e-20)
        assert got == (asym, miss)
        assert miss == 0
        if flux_sym:
            assert asym == 0
        else:
            assert asym > 0
