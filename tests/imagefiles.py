"""PNG and BMP file writers for the front end's image-list tests, written from
the PNG specification (filters: section 9, Adam7 passes: section 8.2) and the
BMP BITMAPINFOHEADER layout. Independent of the decoder under test."""
import struct
import zlib

import numpy as np

# Adam7 passes: (x0, y0, dx, dy)
ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]
CHANNELS = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}


def _filter_rows(img, filters, row0):
    """Filtered scanlines of one (sub-)image img[h, w, ch] (uint8): filter type
    filters[(row0 + y) % len(filters)] on row y."""
    h, w, ch = img.shape
    raw = bytearray()
    prev = np.zeros(w * ch, np.int32)
    for y in range(h):
        cur = img[y].reshape(-1).astype(np.int32)
        f = filters[(row0 + y) % len(filters)]
        left = np.concatenate([np.zeros(ch, np.int32), cur[:-ch]])
        upleft = np.concatenate([np.zeros(ch, np.int32), prev[:-ch]])
        if f == 0:
            out = cur
        elif f == 1:
            out = cur - left
        elif f == 2:
            out = cur - prev
        elif f == 3:
            out = cur - (left + prev) // 2
        else:
            p = left + prev - upleft
            pa, pb, pc = np.abs(p - left), np.abs(p - prev), np.abs(p - upleft)
            pred = np.where((pa <= pb) & (pa <= pc), left, np.where(pb <= pc, prev, upleft))
            out = cur - pred
        raw.append(f)
        raw += bytes((out & 0xFF).astype(np.uint8))
        prev = cur
    return raw


def png_bytes(samples, ctype=2, filters=(0, 1, 2, 3, 4), interlace=False, palette=None, idat_split=40):
    """8-bit PNG of samples[h, w, ch] (ch from the colour type; for type 3 the
    samples are palette indices and `palette` is a list of (r, g, b)). Rows
    cycle through `filters`; the compressed data is split over two IDAT
    chunks at byte `idat_split`."""
    samples = np.asarray(samples, np.uint8)
    if samples.ndim == 2:
        samples = samples[:, :, None]
    h, w, ch = samples.shape
    assert ch == CHANNELS[ctype]
    if interlace:
        raw = bytearray()
        row0 = 0
        for x0, y0, dx, dy in ADAM7:
            sub = samples[y0::dy, x0::dx]
            if sub.shape[0] and sub.shape[1]:
                raw += _filter_rows(sub, filters, row0)
                row0 += sub.shape[0]
    else:
        raw = _filter_rows(samples, filters, 0)

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)
    z = zlib.compress(bytes(raw), 6)
    out = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, ctype, 0, 0, 1 if interlace else 0))
    if ctype == 3:
        out += chunk(b"PLTE", bytes(v for rgb in palette for v in rgb))
    return out + chunk(b"IDAT", z[:idat_split]) + chunk(b"IDAT", z[idat_split:]) + chunk(b"IEND", b"")


def bmp_bytes(rgb, top_down):
    """24-bit BI_RGB BMP of rgb[h, w, 3]; rows bottom-up unless top_down (negative height)."""
    h, w, _ = rgb.shape
    line = (w * 3 + 3) // 4 * 4
    rows = []
    for r in range(h):
        y = r if top_down else h - 1 - r
        row = rgb[y][:, ::-1].reshape(-1).tobytes()  # B, G, R
        rows.append(row + bytes(line - len(row)))
    body = b"".join(rows)
    info = struct.pack("<IiiHHIIiiII", 40, w, -h if top_down else h, 1, 24, 0, len(body), 2835, 2835, 0, 0)
    return struct.pack("<HIHHI", 0x4D42, 54 + len(body), 0, 0, 54) + info + body
