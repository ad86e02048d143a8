"""Read-only subset of the HDF5 file format, enough for Conduit-Blueprint files written by Conduit's
relay (the format dray's BlueprintReader loads through Conduit, io/blueprint_reader.cpp): superblock
versions 2 and 3, version-2 object headers with continuation chunks, groups whose links are stored in
the object header (compact link storage), and datasets with compact or contiguous layout, no filters, of
little-endian integers, IEEE floats or fixed-length strings.  Anything else raises ValueError naming
the unsupported feature.  Written from the published HDF5 File Format Specification (version 3.0).
"""
from __future__ import annotations

import struct

import numpy as np

_UNDEF = 0xFFFFFFFFFFFFFFFF


class H5File:
    """An HDF5 file read into memory; nodes are addressed by '/'-separated paths."""

    def __init__(self, path: str):
        with open(path, "rb") as fh:
            self.buf = fh.read()
        b = self.buf
        if b[:8] != b"\x89HDF\r\n\x1a\n":
            raise ValueError(f"{path}: not an HDF5 file")
        ver = b[8]
        if ver not in (2, 3):
            raise ValueError(f"{path}: superblock version {ver} is not supported (2 and 3 are)")
        if b[9] != 8 or b[10] != 8:
            raise ValueError(f"{path}: offset/length sizes {b[9]}/{b[10]} (8/8 supported)")
        self.root = struct.unpack_from("<Q", b, 36)[0]
        self._cache = {}

    # ---- object headers -------------------------------------------------------------------------
    def _messages(self, addr: int):
        """(type, data) of every message of the version-2 object header at addr, continuations followed."""
        if addr in self._cache:
            return self._cache[addr]
        b = self.buf
        if b[addr:addr + 4] != b"OHDR":
            raise ValueError(f"object header at {addr:#x}: only version-2 headers are supported")
        flags = b[addr + 5]
        p = addr + 6
        if flags & 0x20:
            p += 16              # access / modification / change / birth times
        if flags & 0x10:
            p += 4               # attribute phase-change values
        width = (1, 2, 4, 8)[flags & 3]
        size0 = int.from_bytes(b[p:p + width], "little")
        p += width
        track_order = bool(flags & 0x04)
        out = []

        def parse(start, end, depth=0):
            if depth > 64:
                raise ValueError(f"object header at {addr:#x}: continuation chain too long")
            q = start
            while q + 4 <= end:
                mtype = b[q]
                msize = struct.unpack_from("<H", b, q + 1)[0]
                q += 4 + (2 if track_order else 0)
                data = b[q:q + msize]
                q += msize
                out.append((mtype, data))
                if mtype == 0x10:   # continuation: address and length of an "OCHK" block
                    caddr, clen = struct.unpack_from("<QQ", data, 0)
                    if b[caddr:caddr + 4] != b"OCHK":
                        raise ValueError(f"continuation block at {caddr:#x} has no OCHK signature")
                    if clen < 8:
                        raise ValueError(f"continuation block at {caddr:#x} is too short")
                    parse(caddr + 4, caddr + clen - 4, depth + 1)   # minus the trailing checksum

        parse(p, p + size0)
        self._cache[addr] = out
        return out

    def _links(self, addr: int) -> dict:
        links = {}
        for mtype, d in self._messages(addr):
            if mtype == 0x02:   # link info: dense storage (fractal heap) is not supported
                fl = d[1]
                q = 2 + (8 if fl & 1 else 0)
                heap = struct.unpack_from("<Q", d, q)[0]
                if heap != _UNDEF:
                    raise ValueError("group with dense link storage (fractal heap) is not supported")
            elif mtype == 0x06:   # link
                fl = d[1]
                q = 2
                ltype = 0
                if fl & 0x08:
                    ltype = d[q]
                    q += 1
                if fl & 0x04:
                    q += 8       # creation order
                if fl & 0x10:
                    q += 1       # name character set
                w = (1, 2, 4, 8)[fl & 3]
                n = int.from_bytes(d[q:q + w], "little")
                q += w
                name = d[q:q + n].decode("utf-8")
                q += n
                if ltype != 0:
                    raise ValueError(f"link '{name}': only hard links are supported")
                links[name] = struct.unpack_from("<Q", d, q)[0]
        return links

    def _addr(self, path: str) -> int:
        addr = self.root
        for part in [p for p in path.split("/") if p]:
            links = self._links(addr)
            if part not in links:
                raise KeyError(path)
            addr = links[part]
        return addr

    # ---- public API -----------------------------------------------------------------------------
    def keys(self, path: str = "/") -> list:
        return list(self._links(self._addr(path)))

    def is_group(self, path: str) -> bool:
        return not any(t == 0x08 for t, _ in self._messages(self._addr(path)))

    def read(self, path: str):
        """A dataset's value: numpy array (squeezed to a scalar for one element) or str."""
        msgs = self._messages(self._addr(path))
        dims, dtype, raw, is_str = None, None, None, False
        for mtype, d in msgs:
            if mtype == 0x01:        # dataspace
                ver, nd = d[0], d[1]
                q = 4 if ver == 2 else 8
                dims = [struct.unpack_from("<Q", d, q + 8 * i)[0] for i in range(nd)]
                if ver == 2 and d[3] == 2:
                    dims = [0]   # null dataspace
            elif mtype == 0x03:      # datatype
                cls, bits0 = d[0] & 0x0F, d[1]
                size = struct.unpack_from("<I", d, 4)[0]
                if bits0 & 1 and cls in (0, 1):
                    raise ValueError(f"{path}: big-endian data is not supported")
                if cls == 0:
                    dtype = np.dtype(f"<{'i' if bits0 & 0x08 else 'u'}{size}")
                elif cls == 1:
                    dtype = np.dtype(f"<f{size}")
                elif cls == 3:
                    is_str, dtype = True, np.dtype(f"S{size}")
                else:
                    raise ValueError(f"{path}: datatype class {cls} is not supported")
            elif mtype == 0x0B:
                raise ValueError(f"{path}: filtered (compressed) datasets are not supported")
            elif mtype == 0x08:      # data layout, version 3
                if d[0] != 3:
                    raise ValueError(f"{path}: data layout version {d[0]} is not supported")
                cls = d[1]
                if cls == 0:
                    n = struct.unpack_from("<H", d, 2)[0]
                    raw = d[4:4 + n]
                elif cls == 1:
                    a, n = struct.unpack_from("<QQ", d, 2)
                    raw = b"" if a == _UNDEF else self.buf[a:a + n]
                else:
                    raise ValueError(f"{path}: chunked layout is not supported")
        if raw is None or dtype is None:
            raise ValueError(f"{path} is not a dataset")
        count = 1
        for x in dims or ():
            count *= int(x)
        if not is_str and count * dtype.itemsize > len(raw):
            raise ValueError(f"{path}: dataset shorter than its dataspace")
        if is_str:
            s = bytes(raw[: count * dtype.itemsize])
            return s.split(b"\x00", 1)[0].decode("utf-8")
        arr = np.frombuffer(bytes(raw[: count * dtype.itemsize]), dtype=dtype).copy()
        return arr.reshape(dims) if len(dims) > 1 else (arr[0] if count == 1 and not dims else arr)
