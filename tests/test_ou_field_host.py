"""The library's host-side driving field (maragpu_ou_field_*) and the driving/cooling
verbs' host logic, against the oracle's restatement of StochasticVectorField3d (itself
pinned bitwise to the reference in test_driving_oracle.py). The field's state after
construction and after advances must match bitwise: it is the reference's serialized
byte string."""
import numpy as np
import pytest

from ctf_amd import native
from ctf_amd.mara import Mara, MaraError


@pytest.mark.parametrize("k1,seed,zeta,P0", [(1, 7, 0.0, 1.0), (2, 12345, 0.2, 0.8), (3, 99, 1.0, 2.5)])
def test_field_matches_the_oracle_bitwise(oracle_lib, k1, seed, zeta, P0):
    f = native.OuField.create(P0, zeta, k1, seed)
    ref = oracle_lib.ou_new(P0, zeta, k1, seed)
    assert f.bytes() == ref
    for dt in (0.01, 0.25, 1e-3):
        f.advance(dt)
        ref = oracle_lib.ou_advance(ref, dt, 1)
        assert f.bytes() == ref
    # load / save round trip, and loading the oracle's bytes
    g = native.OuField.load(ref)
    assert g.bytes() == ref
    g.advance(0.1)
    f.advance(0.1)
    assert g.bytes() == f.bytes()


def test_load_rejects_wrong_sizes():
    f = native.OuField.create(1.0, 0.1, 2, 3).bytes()
    with pytest.raises(RuntimeError):
        native.OuField.load(f[:-1])
    with pytest.raises(RuntimeError):
        native.OuField.load(f + b"\0")


def test_driving_and_cooling_verbs_host_logic(oracle_lib):
    Mara.start()
    field = Mara.new_ou_field(3, 0.8, 0.2, 3, 12345)       # drvtrb's field (ctf-core/drvtrb.lua:731)
    assert field == oracle_lib.ou_new(0.8, 0.2, 3, 12345)
    with pytest.raises(MaraError, match="3-d"):
        Mara.new_ou_field(2, 0.8, 0.2, 3, 12345)
    with pytest.raises(MaraError, match="need a domain"):
        Mara.set_driving(field)
    assert Mara.driving.Serialize() is None and Mara.driving.Power() == 0.0
    Mara.set_cooling("X9", 1.0, 1.0)                        # unknown keys leave cooling as it was
    assert Mara.cooling_rate(0.1) == 0.0
    Mara.set_fluid("rmhd")
    Mara.set_domain((0, 0, 0), (1, 1, 1), (4, 4, 4), 8, 3)
    with pytest.raises(MaraError, match="not a serialized"):
        Mara.set_driving(field[:100])
    Mara.set_driving(field)
    assert Mara.driving.Serialize() == field
    Mara.driving.Advance(0.05)
    assert Mara.driving.Serialize() == oracle_lib.ou_advance(field, 0.05, 1)
    Mara.close()
