"""Kernel parameters: the reference's -D macro set as a run-time ``SlmParams``.

The reference renders every kernel constant as a compile-time macro string
``-D NAME=<str(value)><u|f>`` (python/streamlines/pocl.py:178-299) from the
``Info`` object (python/streamlines/useful.py:117-268).  Here
:func:`render_macros` restates that rendering (same names, same numpy-1.x
value types, same string forms) and :func:`params_from_macros` parses the
strings back exactly as a C compiler reads the literals (``strtof`` for
``f``-suffixed floats, modulo-2^32 for ``u`` integers such as ``-1u``).
:func:`kernel_params` chains both, so the struct handed to the HIP kernels
holds bit-for-bit the constants the OpenCL compiler saw.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

ABI_VERSION = 2

_FIELD_NAMES = [
    ("abi_version", ctypes.c_uint32),
    ("nx_padded", ctypes.c_uint32),
    ("ny_padded", ctypes.c_uint32),
    ("nxy_padded", ctypes.c_uint32),
    ("pad_width", ctypes.c_uint32),
    ("pad_width_pp5", ctypes.c_float),
    ("nxf_mp5", ctypes.c_float),
    ("nyf_mp5", ctypes.c_float),
    ("grid_scale", ctypes.c_float),
    ("combo_factor", ctypes.c_float),
    ("dt_max", ctypes.c_float),
    ("adjusted_max_error", ctypes.c_float),
    ("integration_halt_threshold", ctypes.c_float),
    ("max_length", ctypes.c_float),
    ("max_n_steps", ctypes.c_uint32),
    ("trajectory_resolution", ctypes.c_float),
    ("subpixel_seed_point_density", ctypes.c_uint32),
    ("subpixel_seed_halfspan", ctypes.c_float),
    ("subpixel_seed_step", ctypes.c_float),
    ("jitter_magnitude", ctypes.c_float),
    ("downup_sign", ctypes.c_float),
    ("n_seed_points", ctypes.c_uint32),
    ("seeds_chunk_offset", ctypes.c_uint32),
    ("fp_contract", ctypes.c_uint32),
    ("pixel_size", ctypes.c_float),
    ("interchannel_max_n_steps", ctypes.c_uint32),
    ("segmentation_threshold", ctypes.c_uint32),
    ("approx_wrap", ctypes.c_uint32),
    ("reserved", ctypes.c_uint32 * 4),
]


class SlmParams(ctypes.Structure):
    """Mirror of ``struct SlmParams`` in include/slm_hip.h (128 bytes)."""

    _fields_ = _FIELD_NAMES

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in _FIELD_NAMES if n != "reserved"}

    def copy(self) -> "SlmParams":
        c = SlmParams()
        ctypes.memmove(ctypes.byref(c), ctypes.byref(self), ctypes.sizeof(SlmParams))
        return c


assert ctypes.sizeof(SlmParams) == 128

class SlmKdeParams(ctypes.Structure):
    """Mirror of ``struct SlmKdeParams`` in include/slm_hip.h (64 bytes)."""
    _fields_ = [("abi_version", ctypes.c_uint32), ("kernel", ctypes.c_uint32),
                ("n_data", ctypes.c_uint32), ("n_hist_bins", ctypes.c_uint32),
                ("n_pdf_points", ctypes.c_uint32), ("n_kdf_part_points_x", ctypes.c_uint32),
                ("logx_min", ctypes.c_float), ("logx_range", ctypes.c_float),
                ("bin_dx", ctypes.c_float), ("pdf_dx", ctypes.c_float),
                ("kdf_width_x", ctypes.c_float), ("fp_contract", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32 * 4)]


assert ctypes.sizeof(SlmKdeParams) == 64


class SlmNnFlip(ctypes.Structure):
    """Mirror of ``struct SlmNnFlip`` in include/slm_hip.h (32 bytes)."""
    _fields_ = [("offset", ctypes.c_uint64), ("taps", ctypes.c_uint32), ("cin", ctypes.c_uint32),
                ("cout", ctypes.c_uint32), ("reserved", ctypes.c_uint32), ("wt", ctypes.c_void_p)]


assert ctypes.sizeof(SlmNnFlip) == 32


class SlmTileOpts(ctypes.Structure):
    """Mirror of ``struct SlmTileOpts`` in include/slm_hip.h (32 bytes)."""
    _fields_ = [("tile", ctypes.c_uint32), ("block", ctypes.c_uint32), ("unit", ctypes.c_uint32),
                ("cap_steps", ctypes.c_uint32), ("halo", ctypes.c_uint32), ("reserved", ctypes.c_uint32 * 3)]


class SlmFieldsJob(ctypes.Structure):
    """Mirror of ``struct SlmFieldsJob`` in include/slm_hip.h (192 bytes)."""
    _fields_ = [("params", SlmParams), ("seeds_xy", ctypes.c_void_p), ("mask", ctypes.c_void_p),
                ("uv", ctypes.c_void_p), ("slc", ctypes.c_void_p), ("slt", ctypes.c_void_p),
                ("reserved", ctypes.c_uint64 * 3)]


assert ctypes.sizeof(SlmFieldsJob) == 192


class SlmBatchOpts(ctypes.Structure):
    """Mirror of ``struct SlmBatchOpts`` in include/slm_hip.h (32 bytes)."""
    _fields_ = [("chunk", ctypes.c_uint32), ("blocks_per_cu", ctypes.c_uint32), ("queue_slices", ctypes.c_uint32),
                ("lds_slots", ctypes.c_uint32), ("refill_batch", ctypes.c_uint32), ("reserved", ctypes.c_uint32 * 3)]


_libc = ctypes.CDLL(None)
_libc.strtof.restype = ctypes.c_float
_libc.strtof.argtypes = [ctypes.c_char_p, ctypes.c_void_p]


def c_float_literal(text: str) -> np.float32:
    """Value of the C literal ``<text>f`` (correctly rounded decimal->float)."""
    return np.float32(_libc.strtof(text.encode(), None))


def c_uint_literal(text: str) -> int:
    """Value of the C literal ``<text>u``; a leading minus wraps modulo 2^32."""
    return int(text) % (1 << 32)


# The macro table of pocl.py:233-281: name -> (suffix, fixed value or None).
# The fixed values are computed by _integration_values(); None means
# "str(getattr(info, name))".
_MACRO_ORDER = [
    ("n_seed_points", "u"), ("downup_sign", "u"), ("integrator_step_factor", "f"),
    ("max_integration_step_error", "f"), ("adjusted_max_error", "f"),
    ("integration_halt_threshold", "f"), ("max_length", "f"), ("pixel_size", "f"),
    ("pad_width", "u"), ("pad_width_pp5", "f"), ("nx_padded", "u"), ("ny_padded", "u"),
    ("nxy_padded", "u"), ("nxf_mp5", "f"), ("nyf_mp5", "f"), ("grid_scale", "f"),
    ("combo_factor", "f"), ("dt_max", "f"), ("max_n_steps", "u"),
    ("trajectory_resolution", "u"), ("seeds_chunk_offset", "u"),
    ("subpixel_seed_point_density", "u"), ("subpixel_seed_halfspan", "f"),
    ("subpixel_seed_step", "f"), ("jitter_magnitude", "f"), ("interchannel_max_n_steps", "u"),
    ("left_flank_addition", "u"), ("is_channel", "u"), ("is_thinchannel", "u"),
    ("is_interchannel", "u"), ("is_channelhead", "u"), ("is_channeltail", "u"),
    ("is_majorconfluence", "u"), ("is_minorconfluence", "u"), ("is_majorinflow", "u"),
    ("is_minorinflow", "u"), ("is_leftflank", "u"), ("is_rightflank", "u"),
    ("is_midslope", "u"), ("is_ridge", "u"), ("was_channelhead", "u"),
    ("is_subsegmenthead", "u"), ("is_loop", "u"), ("is_blockage", "u"),
    ("do_measure_hsl_from_ridges", "flag"), ("debug", "flag"), ("verbose", "flag"),
]


def _fmt(value) -> str:
    """``'{0}'.format(value)`` as the reference did: shortest round-trip repr."""
    if isinstance(value, (np.floating, float)):
        return str(value.item() if isinstance(value, np.float64) else value)
    if isinstance(value, (np.integer, int, bool, np.bool_)):
        return str(int(value))
    return str(value)


def _integration_values(info, downup_sign) -> dict:
    """Fixed macro values of pocl.py:225-252, with numpy-1.x scalar promotion:
    np.float32 op Python float -> float64 (then str of a float64)."""
    pad = np.uint32(info.pad_width)
    nxp = np.uint32(info.nx_padded)
    nyp = np.uint32(info.ny_padded)
    nxf = np.float32(info.nx)
    nyf = np.float32(info.ny)
    grid_scale = np.float32(np.sqrt(np.float32(nxf * nyf)))
    combo_factor = np.float32(grid_scale * np.float32(info.integrator_step_factor))
    dt_max = np.float32(min(min(1.0 / float(nxf), 1.0 / float(nyf)), 0.1))
    return {
        "downup_sign": int(np.int8(downup_sign)),
        "pad_width": pad,
        "pad_width_pp5": float(np.float32(pad)) + 0.5,
        "nx_padded": nxp,
        "ny_padded": nyp,
        "nxy_padded": np.uint32(int(nxp) * int(nyp)),
        "nxf_mp5": float(nxf) - 0.5,
        "nyf_mp5": float(nyf) - 0.5,
        "grid_scale": grid_scale,
        "combo_factor": float(combo_factor) * float(downup_sign),
        "dt_max": dt_max,
    }


def render_macros(info, kernel_def: str, downup_sign=1) -> list:
    """The ``-D`` option list of pocl.set_compile_options (pocl.py:178-299),
    integration mode.  A false ``flag`` entry re-appends the previous option,
    as the reference's loop does (its ``list_item`` is left unchanged)."""
    fixed = _integration_values(info, downup_sign)
    order = list(_MACRO_ORDER)
    if getattr(info, "segmentation_threshold", 0) > 0:
        order.append(("segmentation_threshold", "u"))
    out = ["-D", "KERNEL_{}".format(kernel_def.upper())]
    item = []
    for name, suffix in order:
        if suffix == "flag":
            if getattr(info, name, False):
                item = ["-D", name.upper()]
        elif name in fixed:
            item = ["-D", "{0}={1}{2}".format(name.upper(), _fmt(fixed[name]), suffix)]
        else:
            item = ["-D", "{0}={1}{2}".format(name.upper(), _fmt(getattr(info, name)), suffix)]
        out += item
    return out


def parse_macros(options: list) -> dict:
    """Macro list -> {NAME: python value} as the C compiler would read them."""
    vals = {}
    for opt in options[1::2]:
        if "=" not in opt:
            vals[opt] = True
            continue
        name, lit = opt.split("=", 1)
        if lit.endswith("u"):
            vals[name] = c_uint_literal(lit[:-1])
        elif lit.endswith("f"):
            vals[name] = c_float_literal(lit[:-1])
        else:
            raise ValueError("unexpected literal " + opt)
    return vals


def params_from_macros(options: list, fp_contract: bool = False, approx_wrap: bool = False) -> SlmParams:
    """SlmParams from the reference's option list.  fp_contract: fuse a*b+c
    where clang's OpenCL default would; off (every operation rounded) is what
    the reference's own recorded runs computed.  approx_wrap: the Euler tail's
    approximate() of an absolute position keeps the low byte of the (char)
    conversion; off truncates it onto the lattice (the int32 value), which is
    what the reference's recorded runs computed (their trajectory statistics
    are reproduced exactly only that way)."""
    m = parse_macros(options)
    p = SlmParams()
    p.abi_version = ABI_VERSION
    for name, ctype in _FIELD_NAMES:
        key = name.upper()
        if key in m:
            setattr(p, name, m[key])
    # DOWNUP_SIGN is used as (float)DOWNUP_SIGN (rungekutta.cl:109, :232)
    p.downup_sign = float(np.float32(m["DOWNUP_SIGN"]))
    p.trajectory_resolution = float(np.float32(m["TRAJECTORY_RESOLUTION"]))
    p.fp_contract = 1 if fp_contract else 0
    p.approx_wrap = 1 if approx_wrap else 0
    return p


def kernel_params(info, downup_sign=1, kernel_def: str = "INTEGRATE_FIELDS",
                  fp_contract: bool = False, approx_wrap: bool = False) -> SlmParams:
    """SlmParams for one kernel launch, via the reference's own macro strings."""
    return params_from_macros(render_macros(info, kernel_def, downup_sign), fp_contract, approx_wrap)


def pad_seed_count(n_seed_points: int, n_work_items: int) -> int:
    """Padded seed count of create_seeds (useful.py:503-505), banker's rounding
    of np.round included."""
    pad_length = int(np.uint32(np.round(n_seed_points / n_work_items + 0.5))) * n_work_items - n_seed_points
    return n_seed_points + pad_length


def _check_float32_math():
    # sqrt of a float32 product must be the float32-rounded sqrt
    assert np.sqrt(np.float32(2.0)).dtype == np.float32
    assert math.isclose(float(c_float_literal("0.1")), float(np.float32(0.1)), rel_tol=0)


_check_float32_math()
