"""Checkpoint files on the host (libmara_h5.so): the HDF5 layout of Mara's Lua drivers,
checked with an independent reader (HDF5's own h5dump) where one is installed, and the JSON
string form of the drivers' json module (dkjson 2.1), checked against its escaping rules
(modules/json.lua:243-304, 507-590) and against Python's own UTF-8 and JSON decoders."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)


@pytest.fixture(scope="module")
def ck():
    from ctf_amd import checkpoint
    if not os.path.exists(checkpoint.H5LIB):
        pytest.skip("libmara_h5.so was not built (no HDF5 installation)")
    try:
        checkpoint.lib()
    except OSError as e:                     # built, but the HDF5 it links is not on this machine
        pytest.skip(f"libmara_h5.so does not load: {e}")
    return checkpoint


def h5dump(*args):
    if H5DUMP is None:
        pytest.skip("h5dump is not installed")
    return subprocess.run([H5DUMP, *args], capture_output=True, text=True, check=True).stdout


def state(N, seed=0):
    from ctf_amd import native
    return native.init_turbulence(N, v_rms=0.3, dB_rms=0.2, seed=seed)


# -- JSON strings ---------------------------------------------------------------------------
KATS = [
    (b"abc", b'"abc"'),
    (b'a"b\\c/d', b'"a\\"b\\\\c/d"'),                       # '/' is not escaped
    (b"\b\f\n\r\t", b'"\\b\\f\\n\\r\\t"'),
    (b"\x00\x01\x1f\x7f", b'"\\u0000\\u0001\\u001f\\u007f"'),
    (b"\xc2\x80\xc2\x9f\xc2\xad", b'"\\u0080\\u009f\\u00ad"'),
    (b"\xc2\xa0", b'"\xc2\xa0"'),                            # not in the second-pass set
    (b"\xd8\x80\xd8\x84\xd8\x85", b'"\\u0600\\u0604\xd8\x85"'),
    (b"\xdc\x8f", b'"\\u070f"'),
    (b"\xe1\x9e\xb4\xe1\x9e\xb5", b'"\\u17b4\\u17b5"'),
    (b"\xe2\x80\x8c\xe2\x80\xa8\xe2\x80\xaf\xe2\x80\xa9", b'"\\u200c\\u2028\\u202f\xe2\x80\xa9"'),
    (b"\xe2\x81\xa0\xe2\x81\xaf", b'"\\u2060\\u206f"'),
    (b"\xef\xbb\xbf\xef\xbf\xb0\xef\xbf\xbf", b'"\\ufeff\\ufff0\\uffff"'),
    (b"\xff\xfe\x80", b'"\xff\xfe\x80"'),                     # other bytes pass through
    (b"\xc2\x0a", b'"\xc2\\n"'),                             # first pass breaks the pair
]


@pytest.mark.parametrize("raw,quoted", KATS)
def test_json_quote_known_answers(ck, raw, quoted):
    assert ck.json_quote(raw) == quoted
    assert ck.json_unquote(quoted) == raw


def test_json_escapes_are_the_code_points(ck):
    # every \\uXXXX of the second pass is the code point of the UTF-8 sequence it replaces,
    # which Python's UTF-8 decoder computes independently
    seqs = [bytes([0xC2, b]) for b in list(range(0x80, 0xA0)) + [0xAD]]
    seqs += [bytes([0xD8, b]) for b in range(0x80, 0x85)] + [b"\xdc\x8f"]
    seqs += [b"\xe1\x9e\xb4", b"\xe1\x9e\xb5", b"\xef\xbb\xbf", b"\xef\xbf\xb0", b"\xef\xbf\xbf"]
    seqs += [bytes([0xE2, 0x80, b]) for b in [0x8C, 0x8D, 0x8E, 0x8F, 0xA8, 0xAF]]
    seqs += [bytes([0xE2, 0x81, b]) for b in range(0xA0, 0xB0)]
    for s in seqs:
        assert ck.json_quote(s) == b'"\\u%04x"' % ord(s.decode("utf-8")), s


def test_json_round_trip_of_binary(ck):
    rng = np.random.default_rng(3)
    for n in (0, 1, 7, 256, 5000):
        data = bytes(range(256)) + rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        q = ck.json_quote(data)
        assert q[:1] == b'"' and q[-1:] == b'"'
        assert ck.json_unquote(q) == data


def test_json_agrees_with_python_on_text(ck):
    import json
    for text in ["plain", "tab\there", "quote\"back\\slash", " ­﻿", "café èà",
                 "\U0001f600 emoji", "ctl\x01\x1f\x7f"]:
        q = ck.json_quote(text.encode())
        assert json.loads(q.decode("utf-8")) == text
        assert ck.json_unquote(json.dumps(text).encode()) == text.encode()


def test_json_unquote_escapes(ck):
    assert ck.json_unquote(b'"\\ud83d\\ude00"') == "\U0001f600".encode()   # surrogate pair
    assert ck.json_unquote(b'"\\ud83dx"') == b"\xed\xa0\xbdx"                # lone high surrogate
    assert ck.json_unquote(b'"\\q\\/"') == b"q/"                              # unknown escape: itself
    assert ck.json_unquote(b'"\\uZZZZ"') == b"uZZZZ"                          # invalid \\u: the letter
    assert ck.json_unquote(b'"ab"tail') == b"ab"
    with pytest.raises(ck.CheckpointError):
        ck.json_unquote(b'"open')


# -- the primitive datasets -------------------------------------------------------------------
def test_prim_layout_seen_by_h5dump(ck, tmp_path):
    N = (3, 4, 5)
    P = state(N, seed=1)
    path = str(tmp_path / "chkpt.h5")
    ck.write_prim(P, path)
    head = h5dump("-H", path)
    assert 'GROUP "prim"' in head
    for name in ck.PRIM_NAMES:
        assert f'DATASET "{name}"' in head
    assert head.count("H5T_IEEE_F64LE") == 8
    assert head.count("DATASPACE  SIMPLE { ( 3, 4, 5 ) / ( 3, 4, 5 ) }") == 8
    # the values, as HDF5's own reader prints them, in (i, j, k) order
    it = P[3:-3, 3:-3, 3:-3]
    for q, name in enumerate(ck.PRIM_NAMES):
        out = h5dump("-d", f"/prim/{name}", "-y", "-w", "0", "-m", "%.17g", path)
        body = out[out.index("DATA {") + 6:out.rindex("}")]
        vals = np.array([float(x) for x in re.findall(r"[-+0-9.eEinfa]+", body)])
        assert np.array_equal(vals, it[..., q].ravel()), name


def test_prim_round_trip_and_ghosts(ck, tmp_path):
    N = (6, 5, 4)
    P = state(N, seed=2)
    path = str(tmp_path / "a.h5")
    ck.write_prim(P, path)
    Q = np.full_like(P, -7.0)
    ck.read_prim(Q, path)
    it = (slice(3, -3),) * 3
    assert np.array_equal(Q[it], P[it])
    ghosts = np.ones(P.shape[:3], bool)
    ghosts[it] = False
    assert (Q[ghosts] == -7.0).all()
    assert ck.File(path).extent("prim/rho") == list(N)
    # update mode replaces the data sets in place
    P2 = P.copy()
    P2[..., 0] *= 2
    ck.write_prim(P2, path, mode=ck.UPDATE)
    ck.read_prim(Q, path)
    assert np.array_equal(Q[it], P2[it])


def test_prim_shape_mismatch_is_an_error(ck, tmp_path):
    path = str(tmp_path / "b.h5")
    ck.write_prim(state((4, 4, 4)), path)
    with pytest.raises(ck.CheckpointError, match="shape"):
        ck.read_prim(np.zeros((11, 10, 10, 8)), path)
    with pytest.raises(ck.CheckpointError):
        ck.read_prim(np.zeros((10, 10, 10, 8)), str(tmp_path / "missing.h5"))


def test_custom_names_and_group(ck, tmp_path):
    path = str(tmp_path / "c.h5")
    f = ck.UnigridDataField((2, 3, 4), Ng=2, dataset_names=["a", "b"])
    A = f.array()
    A[...] = np.arange(A.size, dtype=float).reshape(A.shape)
    f.write(path, "fields")
    g = ck.UnigridDataField((2, 3, 4), Ng=2, dataset_names=["a", "b"])
    g.read(path, "fields")
    it = (slice(2, -2),) * 3
    assert np.array_equal(g.array()[it], A[it])
    assert ck.File(path).keys("fields") == ["a", "b"]


# -- scalar datasets and drvtrb's checkpoints --------------------------------------------------
def test_scalar_datasets(ck, tmp_path):
    path = str(tmp_path / "d.h5")
    ck.write_prim(state((2, 2, 2)), path)
    f = ck.File(path)
    f["version"] = "abc123"
    f["status/Iteration"] = 12
    f["status/CurrentTime"] = 0.25
    f["status/Name"] = "run"
    f["empty"] = ""
    assert f["version"] == b"abc123"
    assert f["status/Iteration"] == 12.0 and f["status/CurrentTime"] == 0.25
    assert f["status/Name"] == b"run"
    assert f["missing"] is None
    f["version"] = "longer text replaces it"                 # clobbered, not appended
    assert f["version"] == b"longer text replaces it"
    assert f.keys("status") == ["CurrentTime", "Iteration", "Name"]
    head = h5dump("-H", "-d", "/version", path)
    assert "STRSIZE 23;" in head and "DATASPACE  SCALAR" in head and "H5T_STR_NULLTERM" in head


def test_drvtrb_checkpoint_round_trip(ck, tmp_path):
    from ctf_amd import native
    N = (4, 4, 4)
    P = state(N, seed=5)
    field = native.OuField.create(P0=0.01, zeta=1.0, k1=2, seed=11)
    field.advance(0.1)
    blob = field.bytes()
    status = {"Iteration": 7, "CurrentTime": 0.125, "Checkpoint": 2, "LastMeasurementTime": 0.1}
    measure_log = {"0.1": {"mean_energies": [1.0, 2.0]}}
    path = str(tmp_path / "chkpt.0002.h5")
    ck.checkpoint_write(P, path, status, measure_log, {"id": "test", "N": 4}, program="-- driver",
                        version="deadbeef", driving=blob)
    Q = np.zeros_like(P)
    st, ml = ck.checkpoint_read(path, Q)
    it = (slice(3, -3),) * 3
    assert np.array_equal(Q[it], P[it])
    assert st == {k: float(v) for k, v in status.items()}
    assert ml == measure_log
    assert ck.read_driving(path) == blob
    assert native.OuField.load(ck.read_driving(path)).bytes() == blob
    f = ck.File(path)
    assert f["version"] == b"deadbeef" and f["program"] == b"-- driver"
    assert set(f.keys()) == {"driving", "measure", "prim", "program", "runargs", "status", "version"}
    # an empty measure log reads back as an empty table (drvtrb.lua:569-573)
    ck.checkpoint_write(P, path, status, [], {})
    assert ck.checkpoint_read(path, Q)[1] == {}
    assert ck.read_driving(path) is None
