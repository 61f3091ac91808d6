"""mlh_riemann_exact_batch_device_noalloc, the allocation-free batch solve
(no counterpart in the reference; for callers that cannot allocate device
memory), against mlh_riemann_exact_batch_device: bitwise the same star
states, flags, iteration counts and fallback counts, on two calls in a row,
the second with a smaller input (the static scratch still holds the first
call's per-block counts), in FP64 and FP32.  Sizes above
MLH_NOALLOC_MAX_FACES are rejected before any device work."""
import numpy as np
import pytest

from meshlesshydro_amd import ics
from meshlesshydro_amd._abi import MLH_ERR_INVALID_ARGUMENT, MLH_FP32, MLH_FP64

pytestmark = pytest.mark.gpu

GAMMA = 5.0 / 3.0
SIZES = ((300_000, 11), (70_001, 12))  # the second call smaller


def _buffers(hip, n, seed):
    rl, ul, pl, rr, ur, pr = ics.riemann_states(n, seed=seed)
    return (hip.DeviceBuffer.from_array(np.stack([rl, ul, pl, rr, ur, pr])),
            hip.DeviceBuffer((3, n), np.float64), hip.DeviceBuffer(n, np.int32),
            hip.DeviceBuffer(n, np.int32))


def _read(bufs):
    _, star, flag, iters = bufs
    out = (star.numpy(), flag.numpy(), iters.numpy())
    for b in bufs:
        b.free()
    return out


@pytest.mark.parametrize("prec", [MLH_FP64, MLH_FP32])
def test_noalloc_matches_allocating_entry_point(hip, prec):
    want = []
    for n, seed in SIZES:
        bufs = _buffers(hip, n, seed)
        fb, _ = hip.riemann_batch_device(GAMMA, *bufs, precision=prec)
        want.append(_read(bufs) + (fb,))
    # two calls in a row, the second smaller
    got = []
    for n, seed in SIZES:
        print(f"[noalloc] precision {prec}, n = {n}", flush=True)
        bufs = _buffers(hip, n, seed)
        fb = hip.riemann_batch_device_noalloc(GAMMA, *bufs, precision=prec)
        got.append(_read(bufs) + (fb,))
    for (n, _), w, g in zip(SIZES, want, got):
        print(f"[noalloc] n = {n}: fallbacks {g[3]} (allocating entry point {w[3]})", flush=True)
        for a, b in zip(w[:3], g[:3]):
            assert np.array_equal(a, b), n
        assert w[3] == g[3], n


def test_noalloc_rejects_oversized_batch(hip):
    from meshlesshydro_amd import lib
    n = (1 << 26) + 1
    st = lib.lib().mlh_riemann_exact_batch_device_noalloc(None, n, GAMMA, None, None, None, None,
                                                          MLH_FP64, 0, 1e-6, None)
    assert st == MLH_ERR_INVALID_ARGUMENT
    assert b"MLH_NOALLOC_MAX_FACES" in lib.lib().mlh_last_error()
