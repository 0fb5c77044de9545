"""A small NetCDF-3 writer and reader (the classic format with 64-bit offsets, "CDF\\x02").

The writer lays out the header once, with every dimension, attribute and variable,
writes fixed-size variables at their offsets, and appends records of the record
variables one at a time, updating the record count in the header after each record,
so a history file is readable (and complete up to its last record) while a run is
still going.  Format: the NetCDF Classic and 64-bit Offset Format specification
(big-endian; names, attribute values and variable data padded to 4 bytes; record
variables interleaved per record, with no padding when only one record variable
exists).
"""
from __future__ import annotations

import struct
from pathlib import Path

import numpy as np

NC_DIMENSION, NC_VARIABLE, NC_ATTRIBUTE = 0x0A, 0x0B, 0x0C
_TYPES = {"i1": 1, "S1": 2, "i2": 3, "i4": 4, "f4": 5, "f8": 6}
_NP = {1: ">i1", 2: "S1", 3: ">i2", 4: ">i4", 5: ">f4", 6: ">f8"}
_SIZE = {1: 1, 2: 1, 3: 2, 4: 4, 5: 4, 6: 8}


def _pad4(n: int) -> int:
    return (n + 3) & ~3


def _name(s: str) -> bytes:
    b = s.encode()
    return struct.pack(">i", len(b)) + b + b"\0" * (_pad4(len(b)) - len(b))


def _attr_value(v):
    """(nc_type, big-endian bytes, element count) of an attribute value."""
    if isinstance(v, (bytes, str)):
        b = v.encode() if isinstance(v, str) else v
        return 2, b, len(b)
    a = np.atleast_1d(np.asarray(v))
    if a.dtype.kind == "f":
        a = a.astype(">f8")
        return 6, a.tobytes(), a.size
    if a.dtype.kind in "iub":
        a = a.astype(">i4")
        return 4, a.tobytes(), a.size
    raise TypeError(f"attribute value of type {a.dtype}")


def _attrs_block(attrs: dict) -> bytes:
    if not attrs:
        return struct.pack(">ii", 0, 0)
    out = [struct.pack(">ii", NC_ATTRIBUTE, len(attrs))]
    for k, v in attrs.items():
        t, b, n = _attr_value(v)
        out.append(_name(k) + struct.pack(">ii", t, n) + b + b"\0" * (_pad4(len(b)) - len(b)))
    return b"".join(out)


class NCWriter:
    """Define dimensions, attributes and variables, then ``begin()`` writes the header;
    ``put`` writes a fixed variable, ``put_record`` one record of a record variable."""

    def __init__(self, path):
        self.path = Path(path)
        self.dims: list[tuple[str, int | None]] = []
        self.gatts: dict = {}
        self.vars: dict[str, dict] = {}
        self.f = None
        self.numrecs = 0

    # ---- definition
    def dim(self, name: str, size: int | None):
        if size is None and any(s is None for _, s in self.dims):
            raise ValueError("only one unlimited dimension")
        self.dims.append((name, None if size is None else int(size)))

    def var(self, name: str, dtype: str, dims: tuple = (), **attrs):
        index = {n: k for k, (n, _) in enumerate(self.dims)}
        ids = [index[d] for d in dims]
        rec = bool(ids) and self.dims[ids[0]][1] is None
        if any(self.dims[i][1] is None for i in ids[1:]):
            raise ValueError(f"{name}: the unlimited dimension must come first")
        shape = tuple(self.dims[i][1] for i in ids[1 if rec else 0:])
        self.vars[name] = dict(type=_TYPES[dtype], ids=ids, rec=rec, shape=shape, attrs=attrs)

    # ---- layout
    def _header(self) -> bytes:
        out = [b"CDF\x02", struct.pack(">i", self.numrecs)]
        if self.dims:
            out.append(struct.pack(">ii", NC_DIMENSION, len(self.dims)))
            for n, s in self.dims:
                out.append(_name(n) + struct.pack(">i", 0 if s is None else s))
        else:
            out.append(struct.pack(">ii", 0, 0))
        out.append(_attrs_block(self.gatts))
        out.append(struct.pack(">ii", NC_VARIABLE, len(self.vars)) if self.vars else struct.pack(">ii", 0, 0))
        for n, v in self.vars.items():
            out.append(_name(n) + struct.pack(">i", len(v["ids"])) + b"".join(struct.pack(">i", i) for i in v["ids"]))
            out.append(_attrs_block(v["attrs"]))
            out.append(struct.pack(">iiq", v["type"], v["vsize"], v["begin"]))
        return b"".join(out)

    def begin(self):
        recvars = [v for v in self.vars.values() if v["rec"]]
        for v in self.vars.values():
            raw = int(np.prod(v["shape"], dtype=np.int64)) * _SIZE[v["type"]]
            v["raw"] = raw
            # a sole record variable is not padded (the format's special case)
            v["vsize"] = raw if (v["rec"] and len(recvars) == 1) else _pad4(raw)
            v["begin"] = 0
        hlen = len(self._header())          # offsets do not change the header length
        off = hlen
        for v in self.vars.values():
            if not v["rec"]:
                v["begin"] = off
                off += v["vsize"]
        self.recsize = 0
        for v in recvars:
            v["begin"] = off + self.recsize
            self.recsize += v["vsize"]
        self.f = open(self.path, "w+b")
        self.f.write(self._header())
        end = off
        self.f.truncate(end)
        return self

    def put(self, name: str, data):
        v = self.vars[name]
        if v["rec"]:
            raise ValueError(f"{name} is a record variable")
        self._write(v, v["begin"], data)

    def put_record(self, name: str, k: int, data):
        v = self.vars[name]
        if not v["rec"]:
            raise ValueError(f"{name} is not a record variable")
        self._write(v, v["begin"] + k * self.recsize, data)
        if k + 1 > self.numrecs:
            self.numrecs = k + 1
            self.f.seek(4)
            self.f.write(struct.pack(">i", self.numrecs))

    def _write(self, v, offset, data):
        a = np.asarray(data, dtype=_NP[v["type"]])
        if a.shape != v["shape"] and a.size != int(np.prod(v["shape"], dtype=np.int64)):
            raise ValueError(f"data of shape {a.shape} for a variable of shape {v['shape']}")
        b = np.ascontiguousarray(a).tobytes()
        self.f.seek(offset)
        self.f.write(b + b"\0" * (v["vsize"] - len(b)))

    def flush(self):
        self.f.flush()

    def close(self):
        if self.f is not None:
            self.f.close()
            self.f = None


class NCReader:
    """Header and variables of a classic or 64-bit offset NetCDF-3 file."""

    def __init__(self, path):
        self.buf = Path(path).read_bytes()
        self.pos = 0
        magic = self.buf[:4]
        if magic[:3] != b"CDF" or magic[3] not in (1, 2):
            raise ValueError(f"{path}: not a NetCDF-3 file")
        self.offset64 = magic[3] == 2
        self.pos = 4
        self.numrecs = self._i()
        self.dims = []
        tag, n = self._i(), self._i()
        for _ in range(n if tag == NC_DIMENSION else 0):
            self.dims.append((self._s(), self._i()))
        self.attrs = self._atts()
        self.vars = {}
        tag, n = self._i(), self._i()
        for _ in range(n if tag == NC_VARIABLE else 0):
            name = self._s()
            ids = [self._i() for _ in range(self._i())]
            attrs = self._atts()
            t, vsize = self._i(), self._i()
            begin = self._q() if self.offset64 else self._i()
            rec = bool(ids) and self.dims[ids[0]][1] == 0
            shape = tuple(self.dims[i][1] for i in ids[1 if rec else 0:])
            self.vars[name] = dict(type=t, vsize=vsize, begin=begin, rec=rec, shape=shape, attrs=attrs,
                                   dims=tuple(self.dims[i][0] for i in ids))
        recvars = [v for v in self.vars.values() if v["rec"]]
        self.recsize = sum(v["vsize"] for v in recvars)

    def _i(self):
        v = struct.unpack_from(">i", self.buf, self.pos)[0]
        self.pos += 4
        return v

    def _q(self):
        v = struct.unpack_from(">q", self.buf, self.pos)[0]
        self.pos += 8
        return v

    def _s(self):
        n = self._i()
        s = self.buf[self.pos:self.pos + n].decode()
        self.pos += _pad4(n)
        return s

    def _atts(self):
        tag, n = self._i(), self._i()
        out = {}
        for _ in range(n if tag == NC_ATTRIBUTE else 0):
            k = self._s()
            t, cnt = self._i(), self._i()
            nb = cnt * _SIZE[t]
            raw = self.buf[self.pos:self.pos + nb]
            self.pos += _pad4(nb)
            out[k] = raw.decode() if t == 2 else np.frombuffer(raw, dtype=_NP[t]).astype(_NP[t][1:])
        return out

    def __getitem__(self, name: str) -> np.ndarray:
        v = self.vars[name]
        dt = np.dtype(_NP[v["type"]])
        count = int(np.prod(v["shape"], dtype=np.int64))
        if not v["rec"]:
            a = np.frombuffer(self.buf, dtype=dt, count=count, offset=v["begin"])
            return a.reshape(v["shape"]).astype(dt.newbyteorder("="))
        recs = [np.frombuffer(self.buf, dtype=dt, count=count, offset=v["begin"] + k * self.recsize)
                for k in range(self.numrecs)]
        out = np.stack(recs) if recs else np.zeros((0,) + v["shape"], dtype=dt)
        return out.reshape((self.numrecs,) + v["shape"]).astype(dt.newbyteorder("="))
