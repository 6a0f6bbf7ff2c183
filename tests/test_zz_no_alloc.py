"""maragpu_measure_noalloc, the measure table without device allocation, against
maragpu_measure on the same committed states: bit-identical results on two calls in a
row on one GPU, the second on a smaller grid (fewer blocks of partials than the first, so
rows left in the shared device array by the first call must not be read)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def state(gpu, N, seed):
    ctx = gpu.Context(gpu.make_config(N=(N, N, N)))
    ctx.upload(gpu.init_turbulence((N, N, N), v_rms=0.4, four_velocity=True, dB_rms=0.3, seed=seed))
    return ctx


def same(a, b):
    assert a.keys() == b.keys()
    for k in a:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k


def test_measure_noalloc_matches_measure(gpu):
    units = (1e8, 2e33, 10.0)
    for N, seed in ((48, 3), (16, 4)):          # 432 blocks of partials, then 16
        print(f"[noalloc] measure table at {N}^3", flush=True)
        ctx = state(gpu, N, seed)
        ref = ctx.measure()
        ref_u = ctx.measure(units)
        got = ctx.measure_noalloc()
        got_u = ctx.measure_noalloc(units)
        same(got, ref)
        same(got_u, ref_u)
        assert ref["max_lorentz_factor"] > 1.0
        ctx.close()
