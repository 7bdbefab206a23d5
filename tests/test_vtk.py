"""The reference's base64 binary point output (.vtp): the restatement of its
encoder pinned against the textbook base64 encoding, and the host writer
(liblbm_host.so) byte for byte against the restatement."""
import base64
import ctypes as ct
import os
import struct

import numpy as np
import pytest

from oracle import vtk_oracle as V

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

FMTS = {"d": np.float64, "f": np.float32, "q": np.int64, "i": np.int32}


def rand_values(fmt, n, rng):
    if fmt in "df":
        v = rng.standard_normal(n) * 10.0 ** rng.integers(-8, 8, n)
        return [float(x) for x in v.astype(FMTS[fmt])]
    return [int(x) for x in rng.integers(-2**31, 2**31 - 1, n)]


@pytest.mark.parametrize("fmt", ["d", "f", "q", "i"])
def test_encoder_restatement_is_textbook_base64(fmt):
    """The bit-level procedure equals standard base64 of the little-endian
    stream [uint64 count*8] + elements, for every length residue (so every
    padding case) -- a wrong element order, byte order, header position or
    bit order fails here."""
    rng = np.random.default_rng(1)
    size = struct.calcsize(fmt)
    for n in range(1, 14):
        vals = rand_values(fmt, n, rng)
        stream = struct.pack("<Q", n * 8) + b"".join(struct.pack("<" + fmt, v) for v in vals)
        std = base64.b64encode(stream)
        assert V.encode_le_header(vals, fmt) == std.rstrip(b"="), (fmt, n)
        assert V.binary(vals, fmt) == std, (fmt, n)
        assert len(V.binary(vals, fmt)) % 4 == 0
        assert base64.b64decode(V.binary(vals, fmt))[8:] == stream[8:]
        # the header counts 8 bytes per element whatever the element size
        assert struct.unpack("<Q", base64.b64decode(V.binary(vals, fmt))[:8])[0] == 8 * n
    with pytest.raises(ValueError):
        V.encode_le_header([], fmt)
    assert size in (4, 8)


def host_lib():
    L = ct.CDLL(os.path.join(ROOT, "lbm_amd", "liblbm_host.so"))
    L.lbm_host_write_points_vtk.argtypes = [ct.c_char_p, ct.c_int32, ct.c_int64, ct.c_void_p, ct.c_int32,
                                            ct.POINTER(ct.c_char_p), ct.POINTER(ct.c_void_p)]
    L.lbm_host_text_round_trip.argtypes = [ct.c_double]
    L.lbm_host_text_round_trip.restype = ct.c_double
    L.lbm_host_last_error.restype = ct.c_char_p
    return L


def host_write(L, base, coords, fields):
    n, dim = coords.shape
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    names = (ct.c_char_p * max(1, len(fields)))(*[f[0].encode() for f in fields])
    keep = [np.ascontiguousarray(f[1], dtype=np.float64) for f in fields]
    vals = (ct.c_void_p * max(1, len(fields)))(*[k.ctypes.data for k in keep])
    return L.lbm_host_write_points_vtk(str(base).encode(), dim, n, coords.ctypes.data, len(fields), names, vals)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_host_writer_bytes_equal_restatement(dim, tmp_path):
    L = host_lib()
    rng = np.random.default_rng(dim)
    for n in (1, 2, 3, 5, 8, 13):  # every residue of the encoded lengths
        coords = rng.uniform(-1, 1, (n, dim)) * 3.7
        names = ["U", "V", "W"][:dim] + ["rho"]
        fields = [(nm, rng.standard_normal(n) * 1e-3) for nm in names]
        base = tmp_path / f"solution_{n}"
        assert host_write(L, base, coords, fields) == 0, L.lbm_host_last_error()
        got = (tmp_path / f"solution_{n}.vtp").read_bytes()
        want = V.write_points(dim, coords.tolist(), [(nm, v.tolist()) for nm, v in fields])
        assert got == want, (dim, n)
        pts, vals, headers = V.read_points(got)
        assert np.array_equal(pts[:, :dim], coords.astype(np.float32))
        assert np.all(pts[:, dim:] == 0)
        for nm, v in fields:
            assert np.array_equal(vals[nm], v)
            assert headers[nm] == 8 * n
        assert headers["Points"] == 8 * 3 * n  # twice the Float32 bytes, as the reference writes
        assert np.array_equal(vals["connectivity"], np.arange(n))


def test_host_writer_rejects_empty(tmp_path):
    L = host_lib()
    assert host_write(L, tmp_path / "x", np.zeros((0, 2)), []) == -1


def test_text_round_trip():
    L = host_lib()
    rng = np.random.default_rng(3)
    vals = list(rng.standard_normal(200) * 10.0 ** rng.integers(-20, 6, 200)) + [0.0, -0.0, 1.0, 1e-16, 5e-16,
                                                                                  0.9999999999999995, 123456.789]
    for v in vals:
        r = L.lbm_host_text_round_trip(v)
        assert r == V.text_round_trip(v) or (r == 0 and V.text_round_trip(v) == 0), v
        # 15 decimals: rounded to 5e-16 absolute, then to the nearest double
        assert abs(r - v) <= 5e-16 + np.spacing(abs(v)), v
