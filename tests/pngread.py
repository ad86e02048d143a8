"""PNG reader for the tests (8-bit RGB/RGBA, non-interlaced, all row filters), independent of libdrx."""
import struct
import zlib

import numpy as np


def _unfilter(raw: np.ndarray, bpp: int) -> np.ndarray:
    """Undo the per-row PNG filters (0 none, 1 sub, 2 up, 3 average, 4 Paeth) of raw [h, 1 + stride]."""
    h, stride = raw.shape[0], raw.shape[1] - 1
    out = np.zeros((h, stride), np.int32)
    prev = np.zeros(stride, np.int32)
    for y in range(h):
        ft, line = int(raw[y, 0]), raw[y, 1:].astype(np.int32)
        if ft == 0:
            cur = line
        elif ft == 2:
            cur = (line + prev) & 0xFF
        else:
            cur = np.zeros(stride, np.int32)
            for x in range(stride):
                a = int(cur[x - bpp]) if x >= bpp else 0
                b = int(prev[x])
                c = int(prev[x - bpp]) if x >= bpp else 0
                if ft == 1:
                    pred = a
                elif ft == 3:
                    pred = (a + b) >> 1
                elif ft == 4:
                    pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
                    pred = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
                else:
                    raise AssertionError(f"filter type {ft}")
                cur[x] = (int(line[x]) + pred) & 0xFF
        out[y] = cur
        prev = cur
    return out.astype(np.uint8)


def read_png(path):
    """Independent PNG reader for 8-bit RGB or RGBA non-interlaced files: checks the signature and every
    chunk's CRC (zlib.crc32), inflates the IDAT stream (zlib) and undoes the row filters -> uint8 [h, w, 4]
    (RGB gets alpha 255, as a 32-bit decode gives)."""
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, ihdr = 8, b"", None
    while pos < len(data):
        (ln,) = struct.unpack(">I", data[pos:pos + 4])
        typ, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + ln]
        (crc,) = struct.unpack(">I", data[pos + 8 + ln:pos + 12 + ln])
        assert crc == zlib.crc32(typ + body) & 0xFFFFFFFF, typ
        if typ == b"IHDR":
            ihdr = struct.unpack(">IIBBBBB", body)
        elif typ == b"IDAT":
            idat += body
        pos += 12 + ln
    assert typ == b"IEND"
    w, h, depth, ctype, comp, filt, inter = ihdr
    assert (depth, comp, filt, inter) == (8, 0, 0, 0) and ctype in (2, 6)
    ch = 4 if ctype == 6 else 3
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + ch * w)
    px = _unfilter(raw, ch).reshape(h, w, ch)
    if ch == 4:
        return px
    return np.concatenate([px, np.full((h, w, 1), 255, np.uint8)], axis=2)
