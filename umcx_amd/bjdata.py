"""Minimal BJData (Binary JData, draft 2) reader/writer for umcx's output files.

umcx writes its volume output as a JNIfTI ``.bnii`` file and its detected photons as a JData
``_detp.jdb`` file, both BJData serialised with counted containers and typed arrays
(src/umcx.cpp:673-705 of fangq/umcx). This module reads those files into Python objects, turning
the JData N-D array form (``[$<type>#[<dims>]<raw little-endian data>``) into numpy arrays, and
can write the same byte layout back (used to cross-check the C++ writer).
"""
from __future__ import annotations

import struct
from typing import Any

import numpy as np

_INT_MARKERS = {
    b"i": ("<b", 1), b"U": ("<B", 1), b"I": ("<h", 2), b"u": ("<H", 2),
    b"l": ("<i", 4), b"m": ("<I", 4), b"L": ("<q", 8), b"M": ("<Q", 8),
}
_FLOAT_MARKERS = {b"h": ("<e", 2), b"d": ("<f", 4), b"D": ("<d", 8)}
_NP_DTYPE = {
    b"i": np.int8, b"U": np.uint8, b"I": np.int16, b"u": np.uint16, b"l": np.int32,
    b"m": np.uint32, b"L": np.int64, b"M": np.uint64, b"h": np.float16, b"d": np.float32,
    b"D": np.float64,
}


class _Reader:
    def __init__(self, buf: bytes):
        self.buf = memoryview(buf)
        self.pos = 0

    def byte(self) -> bytes:
        b = bytes(self.buf[self.pos:self.pos + 1])
        self.pos += 1
        return b

    def peek(self) -> bytes:
        return bytes(self.buf[self.pos:self.pos + 1])

    def number(self, marker: bytes):
        if marker in _INT_MARKERS:
            fmt, n = _INT_MARKERS[marker]
        elif marker in _FLOAT_MARKERS:
            fmt, n = _FLOAT_MARKERS[marker]
        else:
            raise ValueError(f"not a numeric marker: {marker!r} at {self.pos}")
        v = struct.unpack_from(fmt, self.buf, self.pos)[0]
        self.pos += n
        return v

    def length(self) -> int:
        return int(self.number(self.byte()))

    def string(self) -> str:
        n = self.length()
        s = bytes(self.buf[self.pos:self.pos + n]).decode("utf-8")
        self.pos += n
        return s

    def value(self, marker: bytes | None = None) -> Any:
        m = marker if marker is not None else self.byte()
        if m in _INT_MARKERS or m in _FLOAT_MARKERS:
            return self.number(m)
        if m == b"S":
            return self.string()
        if m == b"C":
            return self.byte().decode("latin-1")
        if m == b"T":
            return True
        if m == b"F":
            return False
        if m == b"Z":
            return None
        if m == b"[":
            return self.array()
        if m == b"{":
            return self.obj()
        raise ValueError(f"unsupported BJData marker {m!r} at {self.pos - 1}")

    def array(self):
        typ = None
        if self.peek() == b"$":
            self.pos += 1
            typ = self.byte()
        if self.peek() == b"#":
            self.pos += 1
            if self.peek() == b"[":
                # N-D array: dims given as an array
                self.pos += 1
                dims = self.array()
                n = int(np.prod(dims)) if len(dims) else 0
                dt = np.dtype(_NP_DTYPE[typ]).newbyteorder("<")
                data = np.frombuffer(self.buf, dtype=dt, count=n, offset=self.pos).copy()
                self.pos += n * dt.itemsize
                return data.reshape(dims)  # row-major, as written by umcx
            count = self.length()
            if typ is not None and typ in _NP_DTYPE:
                dt = np.dtype(_NP_DTYPE[typ]).newbyteorder("<")
                data = np.frombuffer(self.buf, dtype=dt, count=count, offset=self.pos).copy()
                self.pos += count * dt.itemsize
                return data.tolist()
            return [self.value(typ) for _ in range(count)]
        out = []
        while self.peek() != b"]":
            out.append(self.value(typ))
        self.pos += 1
        return out

    def obj(self):
        typ = None
        if self.peek() == b"$":
            self.pos += 1
            typ = self.byte()
        out = {}
        if self.peek() == b"#":
            self.pos += 1
            count = self.length()
            for _ in range(count):
                k = self.string()
                out[k] = self.value(typ)
            return out
        while self.peek() != b"}":
            k = self.string()
            out[k] = self.value(typ)
        self.pos += 1
        return out


def loads(buf: bytes) -> Any:
    """Decode one BJData value; JData N-D arrays come back as numpy arrays."""
    return _Reader(buf).value()


def load(path: str) -> Any:
    with open(path, "rb") as f:
        return loads(f.read())


def load_bnii(path: str) -> np.ndarray:
    """Read a umcx ``.bnii`` volume; returns an array indexed [x, y, z, t] (fortran-style view
    of the memory order [t][x][y][z] that umcx writes)."""
    d = load(path)
    data = d["NIFTIData"]
    dims = d["NIFTIHeader"]["Dim"]
    flat = np.asarray(data).reshape(-1)
    nx, ny, nz, nt = (int(v) for v in dims)
    return flat.reshape(nt, nx, ny, nz).transpose(1, 2, 3, 0)


def load_detp(path: str) -> tuple[dict, np.ndarray]:
    """Read a umcx ``_detp.jdb`` file; returns (Info dict, rows [saved, columns])."""
    d = load(path)["MCXData"]
    raw = np.asarray(d["PhotonRawData"], dtype=np.float32)
    info = d["Info"]
    return info, raw.reshape(-1, int(info["ColumnNum"])) if raw.size else raw.reshape(0, int(info["ColumnNum"]))
