"""CPU restatement of the reference's base64 binary VTK point writer (TEST
INFRASTRUCTURE ONLY; pure-Python loops, small point sets).

encode_le_header follows the reference's bit-level procedure step by step
(include/common/util/base64.h:217-270): byte-swap every element and the
header value, lay the bits into one array from the end (pad bits first,
then the elements in reverse order, then the header), read it in 6-bit
groups and write the characters in reverse.  pad follows writeBinary
(src/common/IO.h:392-400) and write_points the layout of
VTK::BINARY::writePoints (src/common/IO.h:414-507) with the XML pieces of
src/common/IO.h:103-208.  The fields pass the fixed-notation text round
trip of toStringVector (include/common/util/string_helper.h:95-105).

Pinned in tests/test_vtk.py against the textbook encoding (Python's base64
module on the little-endian byte stream header + data), which the
bit-level procedure must reproduce for every length and element type.
"""
import math
import struct

TABLE = b"ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/"


def encode_le_header(values, fmt):
    """base64::encodeLE_header<T, GInt>: fmt is the struct code of T."""
    size = struct.calcsize(fmt)
    length = len(values)
    if length == 0:
        raise ValueError("Invalid call to encodeLE() with length = 0")
    actual_bits = size * 8 * length + 64
    num_chars = math.ceil(actual_bits / 6.0)
    mem = [0] * (num_chars * 6)
    # header value: length * BYTE_SIZE (base64.h:235), byte-swapped
    swapped_length = int.from_bytes(struct.pack("<q", length * 8), "big")
    swapped = [struct.pack("<" + fmt, v)[::-1] for v in values]
    count = num_chars * 6 - actual_bits
    for i in range(length):
        elem = swapped[length - i - 1]
        for byte in range(size):
            for bit in range(8):
                mem[count] = (elem[byte] >> bit) & 1
                count += 1
    for bit in range(64):
        mem[count] = (swapped_length >> bit) & 1
        count += 1
    out = bytearray(num_chars)
    for i in range(num_chars):
        num = sum(mem[i * 6 + k] << k for k in range(6))
        out[num_chars - i - 1] = TABLE[num]
    return bytes(out)


def pad(encoded, size, length):
    """writeBinary's padding (IO.h:392-400); the reference's table has no
    entry for padding 4 (no padding due), taken here as none."""
    number_chars = math.ceil((8 + size * length) * 8 / 6.0)
    padding = 4 - number_chars % 4
    return encoded + (b"=" * padding if padding < 4 else b"")


def binary(values, fmt):
    return pad(encode_le_header(values, fmt), struct.calcsize(fmt), len(values))


def text_round_trip(v):
    return float("%.15f" % v)


def write_points(dim, coords, fields):
    """Bytes of the .vtp file: coords [n][dim], fields [(name, values)]."""
    n = len(coords)
    out = [b'<VTKFile type="PolyData" version="1.0" byte_order="LittleEndian" header_type="UInt64">\n'
           b"  <PolyData>\n"]
    out.append(('<Piece NumberOfPoints="%d" NumberOfVerts="1" NumberOfLines="0" NumberOfStrips="0" '
                'NumberOfPolys="0" > \n' % n).encode())
    out.append(b'<Points>\n<DataArray type="Float32" Name="Points" NumberOfComponents="3" format="binary"> \n')
    flat = []
    for c in coords:
        flat += [struct.unpack("<f", struct.pack("<f", float(c[i])))[0] for i in range(dim)] + [0.0] * (3 - dim)
    out.append(binary(flat, "f"))
    out.append(b"\n        </DataArray>\n      </Points>\n")
    out.append(b'      <Verts>\n        <DataArray type="Int64" Name="connectivity" format="binary"> \n')
    out.append(binary(list(range(n)), "q"))
    out.append(b"\n")
    out.append(b"        </DataArray> \n")
    out.append(b'        <DataArray type="Int64" Name="offsets" format="ascii"> \n')
    out.append(b"%d\n" % n)
    out.append(b"        </DataArray> \n")
    out.append(b"        </Verts> \n")
    out.append(b"      <PointData> \n")
    for name, vals in fields:
        out.append(('<DataArray type="Float64" Name="%s" format="binary">\n' % name).encode())
        out.append(binary([float(v) for v in vals], "d"))
        out.append(b"\n        </DataArray> \n")
    out.append(b"      </PointData> \n")
    out.append(b"    </Piece>\n  </PolyData>\n</VTKFile> \n")
    return b"".join(out)


def read_points(data):
    """Decode a .vtp written in this layout: (points [n][3] float32 values,
    {name: float64 values}, header values per array)."""
    import base64
    import re
    import numpy as np
    text = data.decode()
    arrays = re.findall(r'<DataArray type="(\w+)" Name="(\w+)"[^>]*format="binary">\s*\n([A-Za-z0-9+/=]+)\n', text)
    res, headers = {}, {}
    for typ, name, b64 in arrays:
        raw = base64.b64decode(b64)
        headers[name] = struct.unpack("<Q", raw[:8])[0]
        dt = {"Float32": "<f4", "Float64": "<f8", "Int64": "<i8", "Int32": "<i4"}[typ]
        res[name] = np.frombuffer(raw[8:], dtype=dt)
    pts = res.pop("Points").reshape(-1, 3)
    return pts, res, headers
