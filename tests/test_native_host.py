"""Host-side checks of the native library that need no GPU: it loads, exports every
function the public headers declare, the synthetic initial data has the stated
properties, and the Lua-facing verbs reject what Mara's binding rejects."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def declared_functions(headers=None):
    """Functions the public headers declare; by default every header but the checkpoint
    one, whose functions live in libmara_h5.so."""
    names = set()
    paths = glob.glob(os.path.join(ROOT, "include", "*.h"))
    if headers is None:
        paths = [h for h in paths if os.path.basename(h) != "mara_checkpoint.h"]
    else:
        paths = [os.path.join(ROOT, "include", h) for h in headers]
    for h in paths:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b((?:maragpu|mara)_[A-Za-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    from ctf_amd import native
    L = native.lib()
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(native.EXPORTS) <= set(names)


def test_checkpoint_library_exports_every_declared_symbol():
    from ctf_amd import checkpoint
    L = checkpoint.lib()
    names = declared_functions(["mara_checkpoint.h"])
    assert len(names) >= 14
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(checkpoint.EXPORTS) == set(names)


def test_strerror_codes():
    from ctf_amd import native
    L = native.lib()
    for code in (0, -1, -2, -3, -4, -5, -6, 7):
        assert L.maragpu_strerror(code)


def test_initial_data_properties():
    from ctf_amd import native
    N = 32
    P = native.init_turbulence((N, N, N), v_rms=0.3, v_kmax=4.0, dB_rms=0.5, A_kmax=4.0, seed=7)
    it = (slice(3, -3),) * 3
    v = P[it][..., 2:5]
    assert np.isclose(np.sqrt((v**2).sum(-1).mean()), 0.3, rtol=1e-12)
    dB = P[it][..., 5:8] - np.array([0, 0, 1.0])
    assert np.isclose(np.sqrt((dB**2).sum(-1).mean()), 0.5, rtol=1e-12)
    assert np.all(P[..., 0] == 1.0) and np.all(P[..., 1] == 1.0)
    # solenoidal velocity (continuum sense): the spectral divergence vanishes
    vk = [np.fft.fftn(v[..., c]) for c in range(3)]
    k = np.fft.fftfreq(N) * N
    KX, KY, KZ = np.meshgrid(k, k, k, indexing="ij")
    div = KX * vk[0] + KY * vk[1] + KZ * vk[2]
    assert np.abs(div).max() < 1e-9 * np.abs(vk[0]).max()
    # power only on 1 <= |k| <= 4
    kk = np.sqrt(KX**2 + KY**2 + KZ**2)
    power = sum(np.abs(x)**2 for x in vk)
    assert power[(kk > 4.0 + 1e-9) | (kk < 1 - 1e-9)].max() < 1e-18 * power.max()
    # discretely solenoidal field for the corner stencil of Mara/mara.cpp:2298-2314
    from oracle.oracle import Lib, Setup
    so = Lib("oracle").solver(Setup(N=(N, N, N)))
    assert so.mean_max_divB(P)[2] < 1e-11
    # ghosts are periodic images; the same seed gives the same field
    from oracle.oracle import synchronize
    assert np.array_equal(P, synchronize(P, 3))
    Q = native.init_turbulence((N, N, N), v_rms=0.3, v_kmax=4.0, dB_rms=0.5, A_kmax=4.0, seed=7)
    assert np.array_equal(P, Q)


def test_four_velocity_option_is_subluminal():
    from ctf_amd import native
    P = native.init_turbulence((24, 24, 24), v_rms=0.9, v_kmax=8.0, four_velocity=True, seed=3)
    v2 = (P[..., 2:5]**2).sum(-1)
    assert v2.max() < 1.0
    u = P[..., 2:5] / np.sqrt(1 - v2)[..., None]
    it = (slice(3, -3),) * 3
    assert np.isclose(np.sqrt((u[it]**2).sum(-1).mean()), 0.9, rtol=1e-12)


def test_lua_verbs_reject_like_mara():
    from ctf_amd.mara import Mara, MaraError
    with pytest.raises(MaraError):
        Mara.set_fluid("rmhd")              # before Mara.start()
    Mara.start()
    with pytest.raises(MaraError, match="no such fluid"):
        Mara.set_fluid("plasma")
    with pytest.raises(MaraError, match="not supported by this library"):
        Mara.set_fluid("euler")
    with pytest.raises(MaraError, match="no such riemann"):
        Mara.set_riemann("roe")
    with pytest.raises(MaraError, match="no such advance"):
        Mara.set_advance("rk5")
    with pytest.raises(MaraError, match="no such extrap"):
        Mara.config_solver({"extrap": "ppm"}, True)
    with pytest.raises(MaraError, match="no such IS"):
        Mara.config_solver({"IS": "xyz"}, True)
    with pytest.raises(MaraError, match="need a godunov operator"):
        Mara.set_plm_theta(1.5)
    with pytest.raises(MaraError, match="no such integration scheme"):
        Mara.set_godunov("plm-ctu")
    Mara.set_godunov("plm-muscl")
    Mara.set_plm_theta(1.5)
    with pytest.raises(MaraError, match="five fields"):
        Mara.config_solver({"fsplit": "marq"}, True)
    Mara.set_godunov("weno-split")
    Mara.config_solver({"fsplit": "llf"}, True)
    with pytest.raises(MaraError, match="no such eigensystem"):
        Mara.config_solver({"eigensystem": "roe"}, True)
    Mara.config_solver({"eigensystem": "rmhd"}, True)
    Mara.config_solver({"eigensystem": "component"}, True)
    Mara.config_solver({"unknown_key": 3.0, "precision": "mixed", "c2p": "plain"}, True)
    with pytest.raises(MaraError, match="need a domain"):
        Mara.advance(np.zeros(8), 0.0)
    Mara.set_fluid("rmhd")
    Mara.set_domain((0, 0, 0), (1, 1, 1), (4, 4, 4), 8, 3)
    with pytest.raises(MaraError):
        Mara.set_domain((0, 0), (1, 1), (4, 4), 8, 3)   # 2-D is not supported
    Mara.close()


def test_create_rejects_axes_shorter_than_the_ghosts():
    """Configuration checks run before any device call, so they hold without a GPU: a grid
    axis with fewer interior zones than ghosts has ghosts that are not one-period images."""
    import ctypes as C
    from ctf_amd import native
    for N in ((2, 8, 8), (8, 2, 8), (8, 8, 1)):
        cfg = native.make_config(N=N, Ng=3)
        h = C.c_void_p()
        assert native.lib().maragpu_create(C.byref(cfg), C.byref(h)) == -1
        assert not h.value
