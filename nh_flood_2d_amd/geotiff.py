"""Minimal GeoTIFF reader and writer for rainfall rasters (numpy and the standard library only).

Covers what the rain input needs: uncompressed strip TIFFs (chunky or planar bands) of 8/16/32/64
bit integer or float samples, the north-up affine transform from ModelPixelScale + one
ModelTiepoint (shifted by half a pixel when the GeoKey directory declares PixelIsPoint, as GDAL
does), the EPSG code from the GeoKey directory, and GDAL's NODATA tag (compared in the band's own
sample type, as GDAL's mask does). A raster georeferenced any other way (ModelTransformation
matrix, several tiepoints, a missing scale or tiepoint) is refused rather than guessed. Cell
centres are mapped to pixels by flooring the inverse affine transform, the way
``rasterio.transform.rowcol`` does by default (used by the reference at
``preprocess/force.py:185``; its NODATA mask is ``src.sample(..., masked=True)`` at :190).
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
PIXEL_SCALE, TIEPOINT, TRANSFORMATION, GEOKEYS, GDAL_NODATA = 33550, 33922, 34264, 34735, 42113
RASTER_TYPE_KEY, PIXEL_IS_POINT = 1025, 2


@dataclass
class Raster:
    bands: np.ndarray        # [count, height, width] in the file's sample type
    west: float              # edge of the first column (after any PixelIsPoint shift)
    north: float             # edge of the first row
    xres: float
    yres: float
    epsg: int | None
    nodata: np.generic | None   # in the band's sample type; None when absent or unrepresentable

    def rowcol(self, x: np.ndarray, y: np.ndarray):
        rows = np.floor((self.north - np.asarray(y, float)) / self.yres).astype(np.int64)
        cols = np.floor((np.asarray(x, float) - self.west) / self.xres).astype(np.int64)
        return rows, cols

    def is_nodata(self, values: np.ndarray) -> np.ndarray:
        """Mask of `values` (taken from `bands`, still in the sample type) equal to NODATA, or
        any NaN when NODATA is NaN."""
        values = np.asarray(values)
        if self.nodata is None:
            return np.zeros(values.shape, dtype=bool)
        if np.issubdtype(values.dtype, np.floating) and np.isnan(self.nodata):
            return np.isnan(values)
        return values == self.nodata


def nodata_in(text: str | None, dtype: np.dtype):
    """GDAL's NODATA text as a value of the band's sample type: rounded to the nearest float of
    a float band; for an integer band only an integral value in range (anything else matches no
    pixel, None)."""
    if text is None or not text.strip():
        return None
    value = float(text)
    dtype = np.dtype(dtype)
    if np.issubdtype(dtype, np.floating):
        return dtype.type(value)
    info = np.iinfo(dtype)
    if not np.isfinite(value) or value != np.floor(value) or not info.min <= value <= info.max:
        return None
    return dtype.type(int(value))


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
    if TRANSFORMATION in tags:
        raise ValueError(f'{path}: ModelTransformationTag georeferencing is not supported')
    if PIXEL_SCALE not in tags or TIEPOINT not in tags:
        raise ValueError(f'{path}: ModelPixelScale and ModelTiepoint tags are required')
    scale, tie = tags[PIXEL_SCALE], tags[TIEPOINT]
    if len(tie) != 6:
        raise ValueError(f'{path}: exactly one ModelTiepoint is supported')
    if len(scale) < 2 or not scale[0] > 0 or not scale[1] > 0:
        raise ValueError(f'{path}: ModelPixelScale must be positive (north-up raster)')
    west = tie[3] - tie[0] * scale[0]
    north = tie[4] + tie[1] * scale[1]
    epsg, raster_type = None, None
    keys = tags.get(GEOKEYS)
    if keys:
        for k in range(keys[3]):
            key, loc, cnt, val = keys[4 + 4 * k: 8 + 4 * k]
            if key in (3072, 2048) and loc == 0:
                epsg = val
            elif key == RASTER_TYPE_KEY and loc == 0:
                raster_type = val
    if raster_type == PIXEL_IS_POINT:
        # the tiepoint names a pixel's centre: the grid's edges lie half a pixel out
        west -= .5 * scale[0]
        north += .5 * scale[1]
    return Raster(np.array(bands, dtype=dtype.newbyteorder('=')), west, north, float(scale[0]),
                  float(scale[1]), epsg, nodata_in(tags.get(GDAL_NODATA), dtype))


def write_geotiff(path, bands, west: float, north: float, xres: float, yres: float, epsg: int,
                  nodata: float | None = None, dtype='<f4', pixel_is_point: bool = False,
                  transformation: list | None = None, georeference: bool = True) -> None:
    """Planar, one strip per band, north-up. `west`/`north` is the tiepoint (a pixel's corner, or
    with `pixel_is_point` its centre); `transformation` adds a ModelTransformationTag and
    `georeference=False` leaves out the scale and tiepoint tags (cases a reader must refuse)."""
    bands = np.asarray(bands, dtype=np.dtype(dtype).newbyteorder('<'))
    if bands.ndim == 2:
        bands = bands[None]
    count, height, width = bands.shape
    kind = {'f': 3, 'i': 2, 'u': 1}[bands.dtype.kind]
    strip = width * height * bands.dtype.itemsize
    tags = [(256, 4, [width]), (257, 4, [height]), (258, 3, [8 * bands.dtype.itemsize] * count),
            (259, 3, [1]), (262, 3, [1]), (273, 4, [0] * count), (277, 3, [count]), (278, 4, [height]),
            (279, 4, [strip] * count), (284, 3, [2]), (339, 3, [kind] * count)]
    if georeference:
        tags += [(PIXEL_SCALE, 12, [xres, yres, 0.]), (TIEPOINT, 12, [0., 0., 0., west, north, 0.])]
    if transformation is not None:
        tags.append((TRANSFORMATION, 12, list(transformation)))
    keys = [1024, 0, 1, 1, RASTER_TYPE_KEY, 0, 1, PIXEL_IS_POINT if pixel_is_point else 1, 3072, 0, 1, epsg]
    tags.append((GEOKEYS, 3, [1, 1, 0, len(keys) // 4] + keys))
    if nodata is not None:
        tags.append((GDAL_NODATA, 2, repr(float(nodata))))
    tags.sort(key=lambda t: t[0])
    base = 8 + 2 + 12 * len(tags) + 4          # header, then the IFD, then out-of-line values
    blobs = []
    for tag, typ, values in tags:
        fmt, _ = _TYPES[typ]
        raw = values.encode('ascii') + b'\0' if typ == 2 else struct.pack('<' + fmt * len(values), *values)
        blobs.append(raw)
    extra_size = sum(len(raw) + len(raw) % 2 for raw in blobs if len(raw) > 4)
    data_start = base + extra_size
    entries, extra = [], bytearray()
    for (tag, typ, values), raw in zip(tags, blobs):
        if tag == 273:
            raw = struct.pack('<' + 'I' * count, *(data_start + k * strip for k in range(count)))
        cnt = len(raw) if typ == 2 else len(values)
        if len(raw) <= 4:
            entries.append(struct.pack('<HHI', tag, typ, cnt) + raw.ljust(4, b'\0'))
        else:
            entries.append(struct.pack('<HHII', tag, typ, cnt, base + len(extra)))
            extra.extend(raw + b'\0' * (len(raw) % 2))
    with open(path, 'wb') as f:
        f.write(b'II' + struct.pack('<HI', 42, 8))
        f.write(struct.pack('<H', len(entries)) + b''.join(entries) + struct.pack('<I', 0))
        f.write(bytes(extra))
        f.write(bands.tobytes())
