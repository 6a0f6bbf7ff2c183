"""Checkpoints of the device state (maragpu_checkpoint_write / _read): the prim/<field>
datasets hold exactly the committed state's interior, a read sets the committed state with
its periodic ghost images, and a run restarted from a checkpoint continues bit for bit."""
import os

import numpy as np
import pytest

from oracle.oracle import synchronize

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ck():
    from ctf_amd import checkpoint
    if not os.path.exists(checkpoint.H5LIB):
        pytest.skip("libmara_h5.so was not built (no HDF5 installation)")
    try:
        checkpoint.lib()
    except OSError as e:
        pytest.skip(f"libmara_h5.so does not load: {e}")
    return checkpoint


def test_device_write_holds_the_committed_interior(gpu, ck, tmp_path):
    N = (12, 10, 8)
    P = gpu.init_turbulence(N, v_rms=0.3, four_velocity=True, dB_rms=0.2, seed=4)
    ctx = gpu.Context(gpu.make_config(N=N, extrap="plm", riemann="hlld", advance="rk2"))
    ctx.upload(P)
    e, _ = ctx.advance(0.0)
    assert e == 0
    e, _ = ctx.advance(ctx.get_timestep(0.4))
    assert e == 0
    D = ctx.download()
    path = str(tmp_path / "dev.h5")
    ck.write_prim(ctx, path)
    Q = np.zeros_like(D)
    ck.read_prim(Q, path)
    it = (slice(3, -3),) * 3
    assert np.array_equal(Q[it], D[it])
    for q in range(8):
        assert np.array_equal(ctx.download_field(q), D[it][..., q])
    ctx.close()


def test_device_read_sets_state_and_ghosts(gpu, ck, tmp_path):
    N = (8, 12, 10)
    P = gpu.init_turbulence(N, v_rms=0.3, dB_rms=0.2, seed=6)
    path = str(tmp_path / "host.h5")
    ck.write_prim(P, path)
    ctx = gpu.Context(gpu.make_config(N=N))
    ck.read_prim(ctx, path)
    assert np.array_equal(ctx.download(), synchronize(P, 3))
    # a file of another shape is refused and leaves the state alone
    other = str(tmp_path / "other.h5")
    ck.write_prim(gpu.init_turbulence((8, 12, 9), seed=1), other)
    with pytest.raises(ck.CheckpointError, match="shape"):
        ck.read_prim(ctx, other)
    assert np.array_equal(ctx.download(), synchronize(P, 3))
    ctx.close()


@pytest.mark.parametrize("godunov", ["plm-split", "plm-muscl"])
def test_restart_continues_bit_for_bit(gpu, ck, tmp_path, godunov):
    N = (16, 16, 16)
    P = gpu.init_turbulence(N, v_rms=0.4, four_velocity=True, dB_rms=0.3, seed=8)
    cfg = dict(N=N, extrap="plm", riemann="hlld", advance="rk2", godunov=godunov)
    a = gpu.Context(gpu.make_config(**cfg))
    a.upload(P)
    dts = [0.0]
    for _ in range(2):
        e, _ = a.advance(dts[-1])
        assert e == 0
        dts.append(a.get_timestep(0.4))
    path = str(tmp_path / "chkpt.0001.h5")
    ck.checkpoint_write(a, path, {"Iteration": 2, "CurrentTime": sum(dts[:-1])}, {}, {"id": "restart"})
    e, _ = a.advance(dts[-1])
    assert e == 0
    b = gpu.Context(gpu.make_config(**cfg))
    status, log = ck.checkpoint_read(path, b)
    assert status["Iteration"] == 2.0 and log == {}
    e, _ = b.advance(dts[-1])
    assert e == 0
    it = (slice(3, -3),) * 3
    assert np.array_equal(b.download()[it], a.download()[it])
    a.close()
    b.close()
