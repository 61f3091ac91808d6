"""CPU checks of the oracle MLP (oracle/orc_mlp.c): the hand-written
reverse-mode gradient against central differences of the FP64 loss, and the
features against the Toro guesses they encode."""
import numpy as np

from meshlesshydro_amd import ics
from meshlesshydro_amd._abi import MLP_NPARAM


def test_gradient_central_differences(oracle_mod):
    rng = np.random.default_rng(0)
    P = (rng.standard_normal(MLP_NPARAM) * 0.3).astype(np.float32)
    X = rng.standard_normal((64, 10)).astype(np.float32)
    T = rng.standard_normal(64).astype(np.float32)
    loss, g = oracle_mod.mlp_loss_grad(P, X, T)
    assert loss > 0
    for q in rng.choice(MLP_NPARAM, 40, replace=False):
        e = 1e-3
        Pp, Pm = P.copy(), P.copy()
        Pp[q] += e
        Pm[q] -= e
        lp, _ = oracle_mod.mlp_loss_grad(Pp, X, T)
        lm, _ = oracle_mod.mlp_loss_grad(Pm, X, T)
        fd = (lp - lm) / (float(Pp[q]) - float(Pm[q]))
        assert abs(fd - g[q]) <= 2e-3 * max(1e-3, abs(g[q])) + 1e-7, (q, fd, g[q])
    # forward: a network with only the output bias returns it
    Z = np.zeros(MLP_NPARAM, np.float32)
    Z[-1] = 0.75
    assert np.allclose(oracle_mod.mlp_forward(Z, X), 0.75)


def test_features_encode_the_guess(oracle_mod):
    g = 5 / 3
    rl, ul, pl, rr, ur, pr = ics.riemann_states(500, seed=9)
    f, pref = oracle_mod.mlp_features(g, rl, ul, pl, rr, ur, pr)
    assert np.allclose(pref, 0.5 * (pl + pr))
    for k in range(0, 500, 37):
        pg = oracle_mod.riemann_guess(g, rl[k], ul[k], pl[k], rr[k], ur[k], pr[k])
        # feature 7 is log(p_guess / p_ref) of the solver's own adaptive guess
        # (the solver floors it at tol * p_ref)
        if pg > 1e-5 * pref[k]:
            assert abs(f[k, 7] - np.log(pg / pref[k])) < 1e-5
    assert np.all(np.abs(f[:, 8] - (g - 1) / (g + 1)) < 1e-7)


def test_training_set_hash_is_splitmix64():
    """The training set's face hash mixer is splitmix64's output function:
    the generator's published first outputs from seed 0 (state advanced by
    0x9E3779B97F4A7C15 before each output)."""
    from meshlesshydro_amd.learned import mix64
    state = np.array([0, 0x9E3779B97F4A7C15, (2 * 0x9E3779B97F4A7C15) % (1 << 64)], np.uint64)
    out = mix64(state)
    assert [int(v) for v in out] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
