"""Minimal GeoTIFF reader and writer for rainfall rasters (numpy and the standard library only).

Covers what the rain input needs: uncompressed strip TIFFs (chunky or planar bands) of 8/16/32/64
bit integer or float samples, the north-up affine transform from ModelPixelScale + ModelTiepoint,
the EPSG code from the GeoKey directory, and GDAL's NODATA tag. Cell centres are mapped to
pixels by flooring the inverse affine transform, the way ``rasterio.transform.rowcol`` does by
default (used by the reference at ``preprocess/force.py:185``).
"""
from __future__ import annotations

import struct
from dataclasses import dataclass
from pathlib import Path

import numpy as np

_TYPES = {1: ('B', 1), 2: ('s', 1), 3: ('H', 2), 4: ('I', 4), 5: ('II', 8), 11: ('f', 4),
          12: ('d', 8), 16: ('Q', 8)}
_SAMPLE = {(1, 8): np.uint8, (1, 16): np.uint16, (1, 32): np.uint32, (2, 8): np.int8,
           (2, 16): np.int16, (2, 32): np.int32, (3, 32): np.float32, (3, 64): np.float64}


@dataclass
class Raster:
    bands: np.ndarray        # [count, height, width]
    west: float
    north: float
    xres: float
    yres: float
    epsg: int | None
    nodata: float | None

    def rowcol(self, x: np.ndarray, y: np.ndarray):
        rows = np.floor((self.north - np.asarray(y, float)) / self.yres).astype(np.int64)
        cols = np.floor((np.asarray(x, float) - self.west) / self.xres).astype(np.int64)
        return rows, cols


def read_geotiff(path) -> Raster:
    data = Path(path).read_bytes()
    order = {b'II': '<', b'MM': '>'}.get(data[:2])
    if order is None or struct.unpack(order + 'H', data[2:4])[0] != 42:
        raise ValueError(f'{path}: not a classic TIFF')
    (ifd,) = struct.unpack(order + 'I', data[4:8])
    (n,) = struct.unpack(order + 'H', data[ifd:ifd + 2])
    tags = {}
    for i in range(n):
        entry = data[ifd + 2 + 12 * i: ifd + 14 + 12 * i]
        tag, typ, count = struct.unpack(order + 'HHI', entry[:8])
        fmt, size = _TYPES[typ]
        total = size * count
        raw = entry[8:8 + total] if total <= 4 else data[struct.unpack(order + 'I', entry[8:12])[0]:][:total]
        if typ == 2:
            tags[tag] = raw[:count].split(b'\0')[0].decode('ascii', 'replace')
        elif typ == 5:
            vals = struct.unpack(order + 'I' * (2 * count), raw)
            tags[tag] = [vals[2 * k] / vals[2 * k + 1] for k in range(count)]
        else:
            tags[tag] = list(struct.unpack(order + fmt * count, raw))
    width, height = tags[256][0], tags[257][0]
    spp = tags.get(277, [1])[0]
    bits = tags.get(258, [8])[0]
    fmt = tags.get(339, [1])[0]
    if tags.get(259, [1])[0] != 1:
        raise ValueError(f'{path}: compressed TIFFs are not supported')
    if 322 in tags:
        raise ValueError(f'{path}: tiled TIFFs are not supported')
    dtype = np.dtype(_SAMPLE[(fmt, bits)]).newbyteorder(order)
    planar = tags.get(284, [1])[0]
    offsets, counts = tags[273], tags[279]
    blob = b''.join(data[o:o + c] for o, c in zip(offsets, counts))
    values = np.frombuffer(blob, dtype=dtype)
    if planar == 2:
        bands = values.reshape(spp, height, width)
    else:
        bands = values.reshape(height, width, spp).transpose(2, 0, 1)
    scale = tags.get(33550, [1., 1., 0.])
    tie = tags.get(33922, [0., 0., 0., 0., 0., 0.])
    west = tie[3] - tie[0] * scale[0]
    north = tie[4] + tie[1] * scale[1]
    epsg = None
    keys = tags.get(34735)
    if keys:
        for k in range(keys[3]):
            key, loc, cnt, val = keys[4 + 4 * k: 8 + 4 * k]
            if key in (3072, 2048) and loc == 0:
                epsg = val
    nodata = tags.get(42113)
    nodata = float(nodata) if nodata not in (None, '') else None
    return Raster(np.array(bands, dtype=np.float64 if bits == 64 else np.float32), west, north,
                  float(scale[0]), float(scale[1]), epsg, nodata)


def write_geotiff(path, bands, west: float, north: float, xres: float, yres: float, epsg: int,
                  nodata: float | None = None) -> None:
    """Float32, planar, one strip per band, north-up (the layout the tests need)."""
    bands = np.asarray(bands, dtype='<f4')
    if bands.ndim == 2:
        bands = bands[None]
    count, height, width = bands.shape
    entries = []
    extra = bytearray()
    header = 8
    n_tags = 14 + (1 if nodata is not None else 0)
    ifd_size = 2 + 12 * n_tags + 4
    base = header + ifd_size

    def put(tag, typ, values):
        fmt, size = _TYPES[typ]
        if typ == 2:
            raw = values.encode('ascii') + b'\0'
            cnt = len(raw)
        else:
            values = list(values)
            raw = struct.pack('<' + fmt * len(values), *values)
            cnt = len(values)
        if len(raw) <= 4:
            entries.append(struct.pack('<HHI', tag, typ, cnt) + raw.ljust(4, b'\0'))
        else:
            off = base + len(extra)
            extra.extend(raw)
            if len(extra) % 2:
                extra.append(0)
            entries.append(struct.pack('<HHII', tag, typ, cnt, off))

    strip = width * height * 4
    data_start_placeholder = []
    put(256, 4, [width])
    put(257, 4, [height])
    put(258, 3, [32] * count)
    put(259, 3, [1])
    put(262, 3, [1])
    data_start_placeholder.append(len(entries))
    put(273, 4, [0] * count)   # patched below
    put(277, 3, [count])
    put(278, 4, [height])
    put(279, 4, [strip] * count)
    put(284, 3, [2])
    put(339, 3, [3] * count)
    put(33550, 12, [xres, yres, 0.])
    put(33922, 12, [0., 0., 0., west, north, 0.])
    put(34735, 3, [1, 1, 0, 3, 1024, 0, 1, 1, 1025, 0, 1, 1, 3072, 0, 1, epsg])
    if nodata is not None:
        put(42113, 2, repr(float(nodata)))
    data_start = base + len(extra)
    offsets = [data_start + k * strip for k in range(count)]
    # rewrite the StripOffsets entry with the real offsets
    extra_offsets = struct.pack('<' + 'I' * count, *offsets)
    i = data_start_placeholder[0]
    if count == 1:
        entries[i] = struct.pack('<HHI', 273, 4, 1) + extra_offsets
    else:
        tag_off = int(struct.unpack('<I', entries[i][8:12])[0]) - base
        extra[tag_off:tag_off + 4 * count] = extra_offsets
    order = sorted(range(len(entries)), key=lambda k: struct.unpack('<H', entries[k][:2])[0])
    ifd = struct.pack('<H', len(entries)) + b''.join(entries[k] for k in order) + struct.pack('<I', 0)
    with open(path, 'wb') as f:
        f.write(b'II' + struct.pack('<HI', 42, 8))
        f.write(ifd)
        f.write(bytes(extra))
        f.write(bands.tobytes())
