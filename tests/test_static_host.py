"""nhf_run_static's argument checks, on a host without a GPU: they come before any device call.

A mesh whose scratch would not fit the static array is refused with a message that names the
need and the maximum; a NULL mesh, parameters or trace are refused too.
"""
import ctypes

import numpy as np

from nh_flood_2d_amd import abi


def _fake_mesh(e_num: int, s_num: int) -> abi.MeshStruct:
    """Every pointer set to a non-NULL address that is never dereferenced (the size check and the
    argument checks run before the library touches the device)."""
    m = abi.MeshStruct()
    m.e_num, m.s_num, m.r_len, m.b_num, m.n_pixels = e_num, s_num, 2 * s_num, 1, 1
    for name, kind in abi.MESH_FIELDS:
        if kind == 'p' and name not in ('pipe_q', 'pipe_volume', 'pipe_correction', 'open_face', 'x', 'y',
                                        'edge_index', 'hang_offset', 'hang_slot', 'grad_cell', 'grad_axes',
                                        'rain_index'):
            setattr(m, name, 4096)
    return m


def test_static_entry_point_refuses_a_mesh_beyond_its_scratch():
    lib = abi.load_hip_library()
    m = _fake_mesh(20_000_001, 40_000_001)
    p = abi.Params(g=9.81, afa=0.5, theta=1.0, min_h=0.02, batch_capacity=32)
    trace = np.zeros((4, abi.TRACE_COLS), np.uint32)
    coeff = np.zeros(2, np.float32)
    rates = np.zeros(1, np.float32)
    rc = lib.nhf_run_static(ctypes.byref(m), ctypes.byref(p), 0, 1 << 40, 0, coeff.ctypes.data, rates.ctypes.data,
                            1, 4, trace.ctypes.data, None, None)
    assert rc < 0
    text = lib.nhf_last_error().decode()
    assert 'static maximum' in text and str(64 << 20) in text


def test_static_entry_point_refuses_null_arguments():
    lib = abi.load_hip_library()
    m = _fake_mesh(100, 200)
    p = abi.Params(g=9.81, afa=0.5, theta=1.0, min_h=0.02, batch_capacity=32)
    trace = np.zeros((4, abi.TRACE_COLS), np.uint32)
    assert lib.nhf_run_static(None, ctypes.byref(p), 0, 1, 0, None, None, 0, 4, trace.ctypes.data, None, None) < 0
    assert lib.nhf_run_static(ctypes.byref(m), None, 0, 1, 0, None, None, 0, 4, trace.ctypes.data, None, None) < 0
    assert lib.nhf_run_static(ctypes.byref(m), ctypes.byref(p), 0, 1, 0, None, None, 0, 4, None, None, None) < 0
    assert b'null argument' in lib.nhf_last_error()
    m.depth = None
    rc = lib.nhf_run_static(ctypes.byref(m), ctypes.byref(p), 0, 1, 0, None, None, 0, 4, trace.ctypes.data, None, None)
    assert rc < 0 and b'depth' in lib.nhf_last_error()
