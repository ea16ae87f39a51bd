"""Readers for the reference's binary field files (test utility).

Binary format (src/util/util.h:191-258, src/field/field.h:1589-1663):
[u8 len]"size" u32 ncomp [u8]"internal" u32 n [u8]"{" f64[n*ncomp] [u8]"}" ...
"""
import struct

import numpy as np


def _str(buf, pos):
    n = buf[pos]
    return buf[pos + 1:pos + 1 + n].decode(), pos + 1 + n


def read_field_bin(path):
    """Return the internal values of a reference .bin field, shape (n, ncomp)."""
    buf = open(path, "rb").read()
    s, pos = _str(buf, 0)
    assert s == "size", s
    ncomp = struct.unpack_from("<I", buf, pos)[0]
    pos += 4
    s, pos = _str(buf, pos)
    assert s == "internal", s
    n = struct.unpack_from("<I", buf, pos)[0]
    pos += 4
    s, pos = _str(buf, pos)
    assert s == "{", s
    vals = np.frombuffer(buf, dtype="<f8", count=n * ncomp, offset=pos).copy()
    return vals.reshape(n, ncomp)
