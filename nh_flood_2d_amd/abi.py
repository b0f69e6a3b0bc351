"""ctypes mirror of include/nhflood.h and the loader of the HIP library."""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_double, c_float, c_int32, c_int64, c_size_t, c_uint32, c_uint64, c_void_p
from pathlib import Path

ABI_VERSION = 12
TRACE_COLS = 9
MAX_BATCH = 64
MAX_SLOTS = 64
DONE, COMPLETED, BAD_CFL, BAD_DT, OVERSHOOT = range(5)
SIDE_EXTERIOR = 4

PACKAGE_DIR = Path(__file__).resolve().parent
HIP_LIBRARY = PACKAGE_DIR / 'libnhflood.so'


class Params(ctypes.Structure):
    _fields_ = [('g', c_float), ('afa', c_float), ('theta', c_float), ('theta_half', c_float),
                ('min_h', c_float), ('min_dt', c_float), ('batch_capacity', c_int32),
                ('reserved', c_int32)]


# (field name, kind) in header order; kind 'i' = int32 scalar, else an array pointer
MESH_FIELDS = [
    ('e_num', 'i'), ('s_num', 'i'), ('r_len', 'i'), ('b_num', 'i'),
    ('n_pixels', 'i'), ('n_grad', 'i'), ('n_hang', 'i'), ('flags', 'i'),
    ('depth', 'p'), ('bed', 'p'), ('rain_time', 'p'), ('capacity', 'p'),
    ('max_depth', 'p'), ('max_speed', 'p'), ('peak_tick', 'p'),
    ('area', 'p'), ('size', 'p'), ('cell_meta', 'p'), ('class_horton', 'p'),
    ('rain_index', 'p'), ('bweight', 'p'),
    ('csr_start', 'p'), ('csr_count', 'p'), ('pipe_q', 'p'), ('open_face', 'p'),
    ('x', 'p'), ('y', 'p'), ('isl', 'p'), ('isl_cell', 'p'),
    ('q', 'p'), ('side_cells', 'p'), ('side_meta', 'p'),
    ('length', 'p'), ('sdc', 'p'), ('manning', 'p'), ('side_key', 'p'),
    ('edge_index', 'p'), ('hang_offset', 'p'), ('hang_slot', 'p'),
    ('grad_cell', 'p'), ('grad_axes', 'p'),
    ('kind', 'p'),
    ('pipe_volume', 'p'), ('pipe_correction', 'p'),
    ('cell_quad', 'p'),
    ('size_base', 'f'), ('reserved_tail', 'i'),
]


class MeshStruct(ctypes.Structure):
    _fields_ = [(name, {'i': c_int32, 'f': ctypes.c_float, 'p': c_void_p}[kind]) for name, kind in MESH_FIELDS]


# nhb_state: per-scenario state of a batch ([entity][slot] device arrays) and the boundary cells
BATCH_STATE_FIELDS = ['depth', 'rain_time', 'capacity', 'max_depth', 'max_speed', 'peak_tick', 'q',
                      'manning', 'class_manning', 'bslot', 'bcell_start']


class BatchState(ctypes.Structure):
    _fields_ = [(name, c_void_p) for name in BATCH_STATE_FIELDS] + [('n_bcells', c_int32), ('reserved', c_int32)]


class GatesStruct(ctypes.Structure):
    """nhf_gates: gate control lists and state (device arrays for the HIP library)."""
    _fields_ = [('count', c_int32), ('n_cells', c_int32), ('n_items', c_int32), ('n_chunks', c_int32),
                ('trigger_level', c_float), ('opening_seconds', c_float), ('open_elevation', c_float),
                ('min_h', c_float), ('formal_tick', c_uint64)] + [
        (name, c_void_p) for name in ('control', 'cells', 'cell_gate', 'pond', 'closed', 'items', 'chunks',
                                      'flags', 'ticks', 'acc')]


class PartStruct(ctypes.Structure):
    """nhf_part: one mesh of a parts context (cell, side, boundary and pixel ranges)."""
    _fields_ = [(name, c_int32) for name in ('cell_begin', 'cell_end', 'side_begin', 'side_end', 'b_base',
                                             'b_count', 'pix_base', 'pix_count')]


MAX_PARTS = 256
PART_ALIGN = 1024


def bind(lib: ctypes.CDLL) -> ctypes.CDLL:
    """Declare every entry point of include/nhflood.h on a loaded library."""
    vp = c_void_p
    sig = {
        'nhf_abi_version': (c_int32, []),
        'nhf_last_error': (ctypes.c_char_p, []),
        'nhf_workspace_bytes': (c_size_t, [POINTER(MeshStruct), POINTER(Params)]),
        'nhf_create': (c_int32, [POINTER(vp), POINTER(MeshStruct), POINTER(Params), vp, c_size_t, vp]),
        'nhf_destroy': (c_int32, [vp]),
        'nhf_reset': (c_int32, [vp, c_uint64]),
        'nhf_set_forcing': (c_int32, [vp, vp, vp, c_int32]),
        'nhf_begin': (c_int32, [vp, c_uint64, c_uint64]),
        'nhf_run_batch': (c_int32, [vp, c_int32, vp]),
        'nhf_launch_batch': (c_int32, [vp, c_int32]),
        'nhf_wait_batch': (c_int32, [vp, vp]),
        'nhf_compute_cfl': (c_int32, [vp]),
        'nhf_flush_budget': (c_int32, [vp, vp]),
        'nhf_limit_capture': (c_int32, [vp, vp, vp, c_int32]),
        'nhf_cfl_state': (c_int32, [vp, vp, vp]),
        'nhf_sync': (c_int32, [vp]),
        'nhf_set_graph': (c_int32, [vp, c_int32]),
        'nhf_set_timing': (c_int32, [vp, c_int32]),
        'nhf_set_tuning': (c_int32, [vp, c_int32, c_int32, c_int32]),
        'nhf_time_kernels': (c_int32, [vp, c_int32, vp]),
        'nhf_settle_extrema': (c_int32, [vp]),
        'nhf_set_rain_state': (c_int32, [vp, c_int32, c_float, vp]),
        'nhf_set_gates': (c_int32, [vp, POINTER(GatesStruct)]),
        'nhf_gate_check': (c_int32, [vp]),
        'nhf_flush_gate_budget': (c_int32, [vp, vp]),
        'nhf_tuning_range': (c_int32, [vp, vp]),
        'nhf_kernel_times': (c_int32, [vp, vp, vp, c_int32, c_int32]),
        'nhf_clock_probe': (c_int32, [vp, vp, vp, vp, c_float, c_int32, vp, vp, vp]),
        'nhf_rasterize': (c_int32, [vp, vp, vp, vp, vp, c_int32, c_float, c_float, c_float, c_float, c_int32,
                                    c_int32, vp, ctypes.c_int64, vp]),
    }
    if hasattr(lib, 'nhf_run_static'):
        sig['nhf_run_static'] = (c_int32, [POINTER(MeshStruct), POINTER(Params), c_uint64, c_uint64, c_uint64,
                                           vp, vp, c_int32, c_int32, vp, vp, vp])
    if hasattr(lib, 'nhb_create'):
        sig.update({
            'nhb_workspace_bytes': (c_size_t, [POINTER(MeshStruct), POINTER(BatchState), POINTER(Params), c_int32]),
            'nhb_create': (c_int32, [POINTER(vp), POINTER(MeshStruct), POINTER(BatchState), POINTER(Params),
                                     c_int32, vp, c_size_t, vp]),
            'nhb_destroy': (c_int32, [vp]),
            'nhb_reset': (c_int32, [vp, c_int32, c_uint64]),
            'nhb_set_forcing': (c_int32, [vp, c_int32, vp, vp, c_int32]),
            'nhb_begin': (c_int32, [vp, c_int32, c_uint64, c_uint64]),
            'nhb_run_batch': (c_int32, [vp, c_int32, vp]),
            'nhb_launch_batch': (c_int32, [vp, c_int32]),
            'nhb_wait_batch': (c_int32, [vp, vp]),
            'nhb_compute_cfl': (c_int32, [vp]),
            'nhb_flush_budget': (c_int32, [vp, c_int32, vp]),
            'nhb_limit_capture': (c_int32, [vp, c_int32, vp, vp, c_int32]),
            'nhb_cfl_state': (c_int32, [vp, c_int32, vp, vp]),
            'nhb_sync': (c_int32, [vp]),
            'nhb_set_graph': (c_int32, [vp, c_int32]),
            'nhb_set_timing': (c_int32, [vp, c_int32]),
            'nhb_set_tuning': (c_int32, [vp, c_int32, c_int32]),
            'nhb_tuning_range': (c_int32, [vp, vp]),
            'nhb_kernel_times': (c_int32, [vp, vp, vp, c_int32, c_int32]),
            'nhb_settle_state': (c_int32, [vp, c_int32]),
            'nhb_set_rain_state': (c_int32, [vp, c_int32, c_int32, c_float, vp]),
        })
    if hasattr(lib, 'nhu_create'):
        sig.update({
            'nhu_workspace_bytes': (c_size_t, [POINTER(MeshStruct), POINTER(Params), POINTER(PartStruct), c_int32]),
            'nhu_create': (c_int32, [POINTER(vp), POINTER(MeshStruct), POINTER(Params), POINTER(PartStruct),
                                     c_int32, vp, c_size_t, vp]),
            'nhu_reset': (c_int32, [vp, c_int32, c_uint64]),
            'nhu_set_forcing': (c_int32, [vp, c_int32, vp, vp, c_int32]),
            'nhu_begin': (c_int32, [vp, c_int32, c_uint64, c_uint64]),
            'nhu_run_batch': (c_int32, [vp, c_int32, vp]),
            'nhu_launch_batch': (c_int32, [vp, c_int32]),
            'nhu_wait_batch': (c_int32, [vp, vp]),
            'nhu_compute_cfl': (c_int32, [vp]),
            'nhu_flush_budget': (c_int32, [vp, c_int32, vp]),
            'nhu_limit_capture': (c_int32, [vp, c_int32, vp, vp, c_int32]),
            'nhu_cfl_state': (c_int32, [vp, c_int32, vp, vp]),
            'nhu_set_rain_state': (c_int32, [vp, c_int32, c_int32, c_float, vp]),
        })
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.nhf_abi_version() != ABI_VERSION:
        raise RuntimeError(f'{lib._name}: ABI version {lib.nhf_abi_version()} != {ABI_VERSION}')
    return lib


EXPORTED = ['nhf_abi_version', 'nhf_last_error', 'nhf_workspace_bytes', 'nhf_create',
            'nhf_destroy', 'nhf_reset', 'nhf_set_forcing', 'nhf_begin', 'nhf_run_batch',
            'nhf_launch_batch', 'nhf_wait_batch',
            'nhf_compute_cfl', 'nhf_flush_budget', 'nhf_limit_capture', 'nhf_cfl_state',
            'nhf_sync', 'nhf_set_graph', 'nhf_set_timing', 'nhf_kernel_times', 'nhf_clock_probe',
            'nhf_set_tuning', 'nhf_tuning_range', 'nhf_time_kernels', 'nhf_settle_extrema',
            'nhf_set_rain_state',
            'nhf_set_gates', 'nhf_gate_check', 'nhf_flush_gate_budget', 'nhf_run_static',
            'nhb_workspace_bytes', 'nhb_create', 'nhb_destroy', 'nhb_reset', 'nhb_set_forcing', 'nhb_begin',
            'nhb_run_batch', 'nhb_launch_batch', 'nhb_wait_batch', 'nhb_compute_cfl', 'nhb_flush_budget', 'nhb_limit_capture', 'nhb_cfl_state',
            'nhb_sync', 'nhb_set_graph', 'nhb_set_timing', 'nhb_kernel_times', 'nhb_set_tuning',
            'nhb_tuning_range', 'nhb_settle_state', 'nhb_set_rain_state',
            'nhu_workspace_bytes', 'nhu_create', 'nhu_reset', 'nhu_set_forcing', 'nhu_begin', 'nhu_run_batch',
            'nhu_launch_batch', 'nhu_wait_batch', 'nhu_compute_cfl', 'nhu_flush_budget', 'nhu_limit_capture',
            'nhu_cfl_state', 'nhu_set_rain_state', 'nhf_rasterize']

_hip_lib = None


def load_hip_library() -> ctypes.CDLL:
    """Load libnhflood.so (the gfx950 kernels). There is no fallback: a missing build raises."""
    global _hip_lib
    if _hip_lib is None:
        if not HIP_LIBRARY.exists():
            raise RuntimeError(f'{HIP_LIBRARY} is not built; run __graft_entry__.build() first')
        _hip_lib = bind(ctypes.CDLL(str(HIP_LIBRARY)))
    return _hip_lib


def check(lib, rc: int, what: str) -> None:
    if rc < 0:
        raise RuntimeError(f'{what} failed ({rc}): {lib.nhf_last_error().decode(errors="replace")}')


__all__ = ['Params', 'MeshStruct', 'MESH_FIELDS', 'GatesStruct', 'PartStruct', 'MAX_PARTS', 'PART_ALIGN', 'BatchState', 'BATCH_STATE_FIELDS', 'MAX_SLOTS', 'bind', 'load_hip_library', 'check',
           'TRACE_COLS', 'MAX_BATCH', 'DONE', 'COMPLETED', 'BAD_CFL', 'BAD_DT', 'OVERSHOOT',
           'c_double', 'c_int64', 'c_uint32']
