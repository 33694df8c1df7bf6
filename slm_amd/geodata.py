"""DTM input and map output on either side of the tracing path: GeoTIFF
reading and writing, the region of interest, and the DTM / basin masks.

Mirrors python/streamlines/geodata.py (Geodata, :33-294), useful.py
(read_geotiff :315-347, write_geotiff :349-388) and the active-mask part of
state.py (:374-458).  The reference reads and writes through GDAL; here a
small TIFF codec of our own handles what those files are: a single band of
uncompressed strips (any byte order; 8/16/32/64-bit unsigned, signed or
float samples) with the GeoTIFF tags carried through.  The geotransform is
built from ModelPixelScale and ModelTiepoint as GDAL builds it for
pixel-is-area rasters; the projection is kept as the GeoKeyDirectory and
GeoAsciiParams tags and written back unchanged.
"""
from __future__ import annotations

import os
import struct
from dataclasses import dataclass, field

import numpy as np

__all__ = ["TiffInfo", "read_tiff", "write_tiff", "read_geotiff", "write_geotiff", "State", "Geodata"]

# TIFF tags used here
_W, _H, _BITS, _COMPRESSION, _PHOTOMETRIC = 256, 257, 258, 259, 262
_STRIP_OFFSETS, _SPP, _ROWS_PER_STRIP, _STRIP_BYTES, _PLANAR, _SAMPLE_FORMAT = 273, 277, 278, 279, 284, 339
_PIXEL_SCALE, _TIEPOINT, _GEOKEYS, _GEO_DOUBLES, _GEO_ASCII, _GDAL_NODATA = 33550, 33922, 34735, 34736, 34737, 42113
_TYPES = {1: "B", 2: "s", 3: "H", 4: "I", 5: "II", 6: "b", 7: "B", 8: "h", 9: "i", 10: "ii", 11: "f", 12: "d",
          16: "Q", 17: "q"}
_SIZES = {1: 1, 2: 1, 3: 2, 4: 4, 5: 8, 6: 1, 7: 1, 8: 2, 9: 4, 10: 8, 11: 4, 12: 8, 16: 8, 17: 8}
_GEO_TAGS = (_GEOKEYS, _GEO_DOUBLES, _GEO_ASCII, _GDAL_NODATA)


@dataclass
class TiffInfo:
    """What read_tiff found: sizes, the GDAL-style geotransform (None without
    geo tags) and the raw GeoTIFF projection / no-data tags."""
    width: int
    height: int
    dtype: np.dtype
    geotransform: tuple | None = None
    geo_tags: dict = field(default_factory=dict)

    def GetGeoTransform(self):          # the reference reads geodata.tiff.GetGeoTransform()
        return self.geotransform


def _read_ifd(buf, bo, off, big):
    """Tag dict {tag: (type, values)} of the IFD at off (classic or BigTIFF)."""
    if big:
        n = struct.unpack_from(bo + "Q", buf, off)[0]
        off += 8
        esize, cfmt, csize = 20, "Q", 8
    else:
        n = struct.unpack_from(bo + "H", buf, off)[0]
        off += 2
        esize, cfmt, csize = 12, "I", 4
    tags = {}
    for k in range(n):
        e = off + k * esize
        tag, typ = struct.unpack_from(bo + "HH", buf, e)
        cnt = struct.unpack_from(bo + cfmt, buf, e + 4)[0]
        if typ not in _TYPES:
            continue
        size = _SIZES[typ] * cnt
        vo = e + 4 + csize
        if size > csize:
            vo = struct.unpack_from(bo + cfmt, buf, vo)[0]
        if typ == 2:
            tags[tag] = (typ, buf[vo:vo + cnt].decode("latin-1"))
        elif typ in (5, 10):
            v = struct.unpack_from(bo + _TYPES[typ][0] * (2 * cnt), buf, vo)
            tags[tag] = (typ, tuple(v[i] / v[i + 1] for i in range(0, 2 * cnt, 2)))
        else:
            tags[tag] = (typ, struct.unpack_from(bo + _TYPES[typ] * cnt, buf, vo))
    return tags


def read_tiff(filename):
    """(array [height, width], TiffInfo) of the first image of a TIFF with one
    band of uncompressed strips."""
    with open(filename, "rb") as f:
        buf = f.read()
    if buf[:2] == b"II":
        bo = "<"
    elif buf[:2] == b"MM":
        bo = ">"
    else:
        raise ValueError("%s: not a TIFF file" % filename)
    magic = struct.unpack_from(bo + "H", buf, 2)[0]
    if magic == 42:
        big, ifd = False, struct.unpack_from(bo + "I", buf, 4)[0]
    elif magic == 43:
        big, ifd = True, struct.unpack_from(bo + "Q", buf, 8)[0]
    else:
        raise ValueError("%s: not a TIFF file" % filename)
    t = _read_ifd(buf, bo, ifd, big)
    val = lambda tag, d=None: t[tag][1] if tag in t else d
    w, h = val(_W)[0], val(_H)[0]
    if val(_COMPRESSION, (1,))[0] != 1:
        raise ValueError("%s: compressed TIFF (compression %d) is not supported" % (filename, val(_COMPRESSION)[0]))
    if val(_SPP, (1,))[0] != 1:
        raise ValueError("%s: only single-band TIFFs are supported" % filename)
    if _STRIP_OFFSETS not in t:
        raise ValueError("%s: tiled TIFF is not supported" % filename)
    bits = val(_BITS, (8,))[0]
    fmt = val(_SAMPLE_FORMAT, (1,))[0]
    kind = {1: "u", 2: "i", 3: "f"}.get(fmt)
    if kind is None:
        raise ValueError("%s: sample format %d is not supported" % (filename, fmt))
    dt = np.dtype(bo + kind + str(bits // 8))
    raw = b"".join(buf[o:o + c] for o, c in zip(val(_STRIP_OFFSETS), val(_STRIP_BYTES)))
    a = np.frombuffer(raw, dtype=dt, count=w * h).reshape(h, w).astype(dt.newbyteorder("="))
    gt = None
    if _PIXEL_SCALE in t and _TIEPOINT in t:
        sx, sy = val(_PIXEL_SCALE)[:2]
        i, j, _, X, Y, _ = val(_TIEPOINT)[:6]
        gt = (X - i * sx, sx, 0.0, Y + j * sy, 0.0, -sy)
    geo = {k: t[k] for k in _GEO_TAGS if k in t}
    return a, TiffInfo(w, h, a.dtype, gt, geo)


def write_tiff(filename, array, geotransform=None, geo_tags=None):
    """Write a 2-D array [height, width] as a little-endian TIFF: one strip,
    uncompressed, with ModelPixelScale / ModelTiepoint from the GDAL-style
    geotransform and the given GeoTIFF tags ({tag: (type, values)})."""
    a = np.ascontiguousarray(array)
    if a.dtype == np.bool_:
        a = a.astype(np.uint8)
    a = a.astype(a.dtype.newbyteorder("<"))
    h, w = a.shape
    fmt = {"u": 1, "i": 2, "f": 3}[a.dtype.kind]
    entries = {_W: (4, (w,)), _H: (4, (h,)), _BITS: (3, (a.dtype.itemsize * 8,)), _COMPRESSION: (3, (1,)),
               _PHOTOMETRIC: (3, (1,)), _SPP: (3, (1,)), _ROWS_PER_STRIP: (4, (h,)),
               _STRIP_BYTES: (4, (a.nbytes,)), _PLANAR: (3, (1,)), _SAMPLE_FORMAT: (3, (fmt,)),
               _STRIP_OFFSETS: (4, (0,))}
    if geotransform is not None:
        x0, sx, _, y0, _, sy = geotransform
        entries[_PIXEL_SCALE] = (12, (float(sx), float(-sy), 0.0))
        entries[_TIEPOINT] = (12, (0.0, 0.0, 0.0, float(x0), float(y0), 0.0))
    for k, v in (geo_tags or {}).items():
        entries[k] = v
    tags = sorted(entries)
    ifd_off = 8
    data_off = ifd_off + 2 + 12 * len(tags) + 4
    extra, ext_bytes = {}, b""
    for tag in tags:                           # values that do not fit in 4 bytes
        typ, v = entries[tag]
        if typ == 2:
            b = v.encode("latin-1")
        else:
            b = struct.pack("<" + _TYPES[typ] * len(v), *v)
        if len(b) > 4:
            extra[tag] = data_off + len(ext_bytes)
            ext_bytes += b + (b"\0" if len(b) % 2 else b"")
    img_off = data_off + len(ext_bytes)
    img_off += (-img_off) % 16
    entries[_STRIP_OFFSETS] = (4, (img_off,))
    out = bytearray(b"II" + struct.pack("<HI", 42, ifd_off) + struct.pack("<H", len(tags)))
    for tag in tags:
        typ, v = entries[tag]
        cnt = len(v.encode("latin-1")) if typ == 2 else len(v)
        out += struct.pack("<HHI", tag, typ, cnt)
        if tag in extra:
            out += struct.pack("<I", extra[tag])
        else:
            b = v.encode("latin-1") if typ == 2 else struct.pack("<" + _TYPES[typ] * len(v), *v)
            out += b + b"\0" * (4 - len(b))
    out += struct.pack("<I", 0)
    out += ext_bytes
    out += b"\0" * (img_off - len(out))
    out += a.tobytes()
    with open(filename, "wb") as f:
        f.write(out)


def read_geotiff(path, filename):
    """useful.py:315-347: (grid float32 [rows, cols], TiffInfo, pixel size);
    ValueError for a missing file or non-square pixels."""
    full = os.path.join(path, filename)
    if not os.path.exists(full):
        raise ValueError('Cannot open GeoTIFF file "{}" for reading'.format(full))
    a, info = read_tiff(full)
    gt = info.geotransform or (0.0, 1.0, 0.0, 0.0, 0.0, -1.0)
    pixel_size = gt[1]
    if not np.isclose(pixel_size, gt[5] * (-1), rtol=1e-3):
        raise ValueError("Pixel x={0} and y={1} dimensions not equal in \"{2}\": cannot handle non-square pixels"
                         .format(pixel_size, (-1) * gt[5], full))
    return a.astype(np.float32), info, pixel_size


def write_geotiff(path, file_name, array, nx, ny, npd, pslice, geodata):
    """useful.py:349-388: write array[pslice], rotated back to the file's
    orientation (flipud of the transpose), with the ROI geotransform and the
    DTM's projection tags.  Booleans are written as uint8."""
    rotated = np.flipud(np.asarray(array)[pslice].T)
    if rotated.shape != (ny, nx):
        raise ValueError("grid %s does not match nx=%d, ny=%d" % (rotated.shape, nx, ny))
    write_tiff(os.path.join(path, file_name), rotated, geodata.roi_geotransform,
               {k: v for k, v in geodata.tiff.geo_tags.items() if k != _GDAL_NODATA})


class State:
    """The active masks of state.py:374-458 and the JSON state keys."""

    def __init__(self, params=None, parameters_path=None):
        for k, v in (params or {}).items():
            setattr(self, k, v)
        self.parameters_path = parameters_path or os.getcwd()
        self.active_masks_dict = {}

    def add_active_mask(self, mask_dict):
        self.active_masks_dict.update(mask_dict)

    def reset_active_masks(self):
        keep = ["dtm", "basin", "uv"]
        self.active_masks_dict = {k: v for k, v in self.active_masks_dict.items() if k in keep}

    def merge_active_masks(self, out=None):
        if not self.active_masks_dict:
            return None
        for idx, m in enumerate(self.active_masks_dict.values()):
            if idx == 0:
                if out is None:
                    out = m.copy()
                else:
                    np.copyto(out, m)
            else:
                out |= m
        return out


class Geodata:
    """Geodata (geodata.py:33-294): read the DTM, cut the region of interest,
    build the padded DTM mask (and the basin mask).  params: the JSON geodata
    keys (dtm_file, data_path, no_data_values, do_clip_roi, roi_bounds_units,
    roi_x/y_bounds[_meters], h_min, flip_ns, pad_width, basins_file, basins,
    do_basin_masking); dtm_path may be given to skip resolving data_path
    against the parameters file's folder."""

    def __init__(self, state, params, dtm_path=None):
        self.state = state
        p = vars(params) if not isinstance(params, dict) else params
        defaults = {"no_data_values": [-99999, 0], "do_clip_roi": False, "roi_bounds_units": "pixels",
                    "roi_x_bounds": [], "roi_y_bounds": [], "roi_x_bounds_meters": [], "roi_y_bounds_meters": [],
                    "h_min": "none", "flip_ns": False, "pad_width": 2, "basins_file": "", "basins": [],
                    "do_basin_masking": False, "data_path": [""]}
        for k, v in {**defaults, **p}.items():
            setattr(self, k, v)
        self.roi_x_bounds = list(self.roi_x_bounds)
        self.roi_y_bounds = list(self.roi_y_bounds)
        if dtm_path is not None:
            self.dtm_path = dtm_path
        self.active_masks_dict = {}

    def do(self):
        self.read_dtm_file()
        self.make_dtm_mask()
        if self.do_basin_masking:
            self.read_basins_file()
            self.make_basins_mask()

    def _path(self):
        if not hasattr(self, "dtm_path"):
            self.dtm_path = os.path.dirname(os.path.realpath(
                os.path.join(os.path.realpath(self.state.parameters_path), *self.data_path, self.dtm_file)))
        return self.dtm_path

    def read_dtm_file(self):
        """geodata.py:65-226."""
        self.dtm_array, self.tiff, self.pixel_size = read_geotiff(self._path(), self.dtm_file)
        gt = self.tiff.GetGeoTransform() or (0.0, self.pixel_size, 0.0, 0.0, 0.0, -self.pixel_size)
        self.x_easting_bottomleft = gt[0]
        self.y_northing_bottomleft = gt[3] + gt[5] * self.dtm_array.shape[0]
        for self.no_data_value in self.no_data_values:
            self.dtm_array[self.dtm_array == self.no_data_value] = np.nan
        rows, cols = self.dtm_array.shape
        if not self.do_clip_roi or (self.roi_x_bounds == [] and self.roi_x_bounds_meters == []):
            self.roi_x_bounds = [0, cols]
            self.roi_x_bounds_meters = [0, cols * self.pixel_size]
        if not self.do_clip_roi or (self.roi_y_bounds == [] and self.roi_y_bounds_meters == []):
            self.roi_y_bounds = [0, rows]
            self.roi_y_bounds_meters = [0, rows * self.pixel_size]
        if self.do_clip_roi and self.roi_bounds_units == "meters":
            self.roi_x_bounds = [int(np.round(self.roi_x_bounds_meters[0] / self.pixel_size)),
                                 int(np.round(self.roi_x_bounds_meters[1] / self.pixel_size))]
            self.roi_y_bounds = [int(np.round(self.roi_y_bounds_meters[0] / self.pixel_size)),
                                 int(np.round(self.roi_y_bounds_meters[1] / self.pixel_size))]
        else:
            self.roi_x_bounds_meters = [self.roi_x_bounds[0] * self.pixel_size,
                                        self.roi_x_bounds[1] * self.pixel_size]
            self.roi_y_bounds_meters = [self.roi_y_bounds[0] * self.pixel_size,
                                        self.roi_y_bounds[1] * self.pixel_size]
        if np.any(np.array(self.roi_x_bounds) != np.clip(np.array(self.roi_x_bounds), 0, cols)):
            raise ValueError("ROI out of bounds in x: " + str(np.array(self.roi_x_bounds)) + " > " + str(cols))
        if np.any(np.array(self.roi_y_bounds) != np.clip(np.array(self.roi_y_bounds), 0, rows)):
            raise ValueError("ROI out of bounds in y: " + str(np.array(self.roi_y_bounds)) + " > " + str(rows))
        x0, x1 = self.roi_x_bounds
        y0, y1 = self.roi_y_bounds
        self.roi_array = np.flipud(self.dtm_array)[y0:y1, x0:x1].T.copy()
        self.x_roi_n_pixel_centers = np.linspace(x0 + 0.5, x1 - 0.5, self.roi_array.shape[0], dtype=np.float32)
        self.y_roi_n_pixel_centers = np.linspace(y0 + 0.5, y1 - 0.5, self.roi_array.shape[1], dtype=np.float32)
        self.roi_nx = len(self.x_roi_n_pixel_centers)
        self.roi_ny = len(self.y_roi_n_pixel_centers)
        self.roi_x_origin = self.x_roi_n_pixel_centers[0]
        self.roi_y_origin = self.y_roi_n_pixel_centers[0]
        self.roi_pixel_size = self.pixel_size
        if self.h_min != "none":
            self.roi_array[np.isnan(self.roi_array)] = self.h_min
        if self.flip_ns:
            self.roi_array = np.fliplr(self.roi_array)
            self.dtm_array = np.fliplr(self.dtm_array)
        self.roi_geotransform = list(gt)
        self.roi_geotransform[0] += x0 * self.pixel_size
        self.roi_geotransform[3] -= rows * self.pixel_size
        self.roi_geotransform[3] += y1 * self.pixel_size

    def read_basins_file(self):
        """geodata.py:228-250."""
        basins, _, self.geotransform = read_geotiff(self._path(), self.basins_file)
        if basins.shape != self.dtm_array.shape:
            raise ValueError("DTM grid and basins grid sizes do not match - DTM: %s, basins=%s"
                             % (str(self.dtm_array.shape), str(basins.shape)))
        x0, x1 = self.roi_x_bounds
        y0, y1 = self.roi_y_bounds
        self.basins_array = np.flipud(basins)[y0:y1, x0:x1].T.copy()
        if self.flip_ns:
            self.basins_array = np.fliplr(self.basins_array)

    def make_dtm_mask(self):
        """geodata.py:252-273: NaN (and h <= h_min) masked, padded with True."""
        m = np.zeros_like(self.roi_array, dtype=np.bool_)
        m[np.isnan(self.roi_array)] = True
        if self.h_min != "none":
            m[self.roi_array <= self.h_min] = True
        pad = int(self.pad_width)
        self.dtm_mask_array = np.pad(m, (pad, pad), "constant", constant_values=(True, True))
        self.state.add_active_mask({"dtm": self.dtm_mask_array})

    def make_basins_mask(self):
        """geodata.py:275-294: everything but the listed basins masked."""
        keep = np.zeros_like(self.basins_array, dtype=np.bool_)
        for basin in self.basins:
            keep[self.basins_array == basin] = True
        pad = int(self.pad_width)
        self.basin_mask_array = np.pad(~keep, (pad, pad), "constant", constant_values=(True, True))
        self.state.add_active_mask({"basin": self.basin_mask_array})
