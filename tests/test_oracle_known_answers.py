"""Pin the CPU oracle with the reference's own known answers (no GPU).

Each test restates one known-answer test of the reference against the oracle:
  clock         tests/test_clock.py:30-164
  budget        tests/test_device_budget.py:40-167
  face average  tests/test_face_flux.py:29-110
  open faces    tests/test_open_boundary.py:64-94
  pressure      tests/test_pressure_gradient.py:154-230
"""
import ctypes
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle
from nh_flood_2d_amd import abi
from nh_flood_2d_amd.clock import seconds_to_ticks
from oracle.budget import ExactBudget

SCALE = 1 << 40


@pytest.fixture(scope='module')
def lib():
    lib = oracle.load('f32')
    vp = ctypes.c_void_p
    lib.nhf_oracle_face_discharge.argtypes = [vp, vp, vp, ctypes.c_int32, ctypes.c_int32, vp]
    lib.nhf_oracle_neighbor_sum.argtypes = [vp, ctypes.c_int32, vp]
    lib.nhf_oracle_theta_update.argtypes = [vp, ctypes.c_int32, ctypes.c_float, ctypes.c_float, vp]
    lib.nhf_oracle_pressure.argtypes = [vp, ctypes.c_float, ctypes.c_int32, vp, vp, vp, vp]
    lib.nhf_oracle_dd_sum.argtypes = [vp, ctypes.c_int64, vp]
    return lib


def probe(lib, now, target, origin, candidate, min_dt=1e-4):
    n = 1
    a = [np.array([v], dtype=np.uint64) for v in (now, target, origin)]
    c = np.array([candidate], dtype=np.float32)
    words = np.zeros(4, np.uint32)
    ticks = np.zeros(2, np.uint64)
    times = np.zeros(2, np.float32)
    assert lib.nhf_clock_probe(a[0].ctypes.data, a[1].ctypes.data, a[2].ctypes.data, c.ctypes.data,
                               np.float32(min_dt), n, words.ctypes.data, ticks.ctypes.data,
                               times.ctypes.data) == 0
    return dict(status=int(words[0]), dt=words[1:2].view(np.float32)[0], cfl_bits=int(words[2]),
                short=int(words[3]), end=int(ticks[0]), local=float(times[0]), end_local=float(times[1]))


# ------------------------------------------------------------------------------ clock

def test_regular_f32_dt_is_encoded_exactly(lib):
    r = probe(lib, 0, seconds_to_ticks(1.), 0, 0.2)
    assert r['status'] == abi.COMPLETED
    assert r['local'] == 0. and abs(r['end_local'] - float(np.float32(0.2))) < 1e-8
    expected = Fraction.from_float(float(np.float32(0.2))) * SCALE
    assert expected.denominator == 1 and r['end'] == expected.numerator
    assert r['dt'] == np.float32(0.2)


def test_short_event_below_min_dt_uses_a_positive_step(lib):
    now = seconds_to_ticks(169199.9999999)
    target = seconds_to_ticks(169200.)
    r = probe(lib, now, target, now, 0.25)
    assert r['status'] == abi.COMPLETED and 0. < float(r['dt']) < 1e-4
    assert r['end'] == 169200 * SCALE
    again = probe(lib, r['end'], target, now, 0.25)
    assert again['status'] == abi.DONE


def test_two_word_carry(lib):
    initial = 1 / 256 - 0.0002
    now = seconds_to_ticks(initial)
    r = probe(lib, now, seconds_to_ticks(initial + 0.0005), now, 0.0005)
    assert r['status'] == abi.COMPLETED
    assert r['end'] >> 32 == 1 and r['end'] & 0xffffffff > 0


@pytest.mark.parametrize('initial, snap', [(169199.7999991, True), (169199.7999989, False)])
def test_snap_threshold(lib, initial, snap):
    now = seconds_to_ticks(initial)
    r = probe(lib, now, seconds_to_ticks(169200.), now, 0.2)
    assert r['status'] == abi.COMPLETED
    assert (r['end'] == 169200 * SCALE) is snap


def test_bad_cfl_stops_before_physics(lib):
    r = probe(lib, 0, seconds_to_ticks(1.), 0, 5e-5)
    assert r['status'] == abi.BAD_CFL and r['end'] == 0


def test_target_reached_gives_done(lib):
    t = seconds_to_ticks(47 * 3600.)
    assert probe(lib, t, t, t, 0.25)['status'] == abi.DONE


def test_47h_end_step_keeps_dt_and_event_tick(lib):
    now = seconds_to_ticks(47 * 3600. - 0.25)
    r = probe(lib, now, seconds_to_ticks(47 * 3600.), now, 0.25)
    assert r['dt'] == np.float32(0.25) and r['end'] == 47 * 3600 * SCALE


def test_host_rejects_out_of_range_clock():
    with pytest.raises(ValueError):
        seconds_to_ticks(-1.)
    with pytest.raises(ValueError):
        seconds_to_ticks(194.181 * 86400.)


def test_sub_microsecond_candidate_rounds_half_even(lib):
    small = np.nextafter(np.float32(2 ** -20), np.float32(1.0))
    exact = Fraction.from_float(float(small)) * SCALE
    assert exact.denominator > 1
    r = probe(lib, 0, seconds_to_ticks(1.), 0, float(small), min_dt=float(small))
    assert r['status'] == abi.COMPLETED
    assert r['end'] == round(exact)
    assert abs(r['end'] / SCALE - float(small)) <= 0.5 / SCALE


# ------------------------------------------------------------------------------ budget

UNIT = 1 << 149


def exact_units(value):
    return int(Fraction.from_float(float(value)) * UNIT)


def exact_units_from_bits(bits):
    bits = int(bits)
    exponent = (bits >> 23) & 255
    fraction = bits & 0x7fffff
    sign = -1 if bits >> 31 else 1
    significand = Fraction(fraction + ((1 << 23) if exponent else 0), 1 << 23)
    power = exponent - 127 if exponent else -126
    return int(sign * significand * Fraction(2) ** power * UNIT)


def inject(budget, values):
    budget.step[...] = np.asarray(values, dtype=np.float32)
    budget.collect()
    budget.step_finished()


def test_budget_cancellation_keeps_small_values():
    budget = ExactBudget(2, 2, 2)
    large = np.float32(1e30)
    inject(budget, [[2 ** 25, large], [1, 1]])
    first = budget.flush()
    assert budget.total_units == [exact_units(np.float32(2 ** 25)) + UNIT, exact_units(large) + UNIT]
    assert first[0] == 2 ** 25 + 1 and first[1] == float(large)
    inject(budget, [[-2 ** 25, -large], [0, 0]])
    np.testing.assert_array_equal(budget.flush(), [1., 1.])
    assert budget.total_units == [UNIT, UNIT] and budget.transfer_count == 2


def test_budget_signed_subnormals():
    budget = ExactBudget(2, 5, 1)
    bits = np.array([[0x00000001, 0x80000001, 0x007fffff, 0x80000000, 0x00800000],
                     [0x80000001, 0x00000001, 0x80800000, 0x00000000, 0x007fffff]], dtype=np.uint32)
    inject(budget, bits.view(np.float32))
    result = budget.flush()
    assert budget.total_units == [0, 0, -1, 0, (1 << 24) - 1]
    np.testing.assert_array_equal(result, [0., 0., -math.ldexp(1., -149), 0.,
                                           math.ldexp(float((1 << 24) - 1), -149)])


def test_budget_rounds_the_complete_integer_once():
    budget = ExactBudget(3, 1, 1)
    inject(budget, [[2 ** 27], [2 ** -26], [2 ** -126]])
    assert budget.flush()[0].hex() == '0x1.0000000000001p+27'


def test_budget_random_patterns_match_fractions():
    rng = np.random.default_rng(48193)
    budget = ExactBudget(7, 4, 40)
    reference = [0] * 4
    for step in range(40):
        bits = rng.integers(0, 1 << 32, size=(7, 4), dtype=np.uint32)
        nonfinite = ((bits >> 23) & 255) == 255
        bits[nonfinite] ^= np.uint32(1 << 23)
        inject(budget, bits.view(np.float32))
        for term in range(4):
            reference[term] += sum(exact_units_from_bits(v) for v in bits[:, term])
        if step in (16, 39):
            result = budget.flush()
            assert budget.total_units == reference
            np.testing.assert_array_equal(result, [math.ldexp(float(u), -149) for u in reference])


def test_budget_max_finite_carries_and_nonfinite_fails():
    budget = ExactBudget(1, 1, 32)
    bits = np.array([[0x7f7fffff]], dtype=np.uint32)
    for _ in range(32):
        inject(budget, bits.view(np.float32))
    budget.flush()
    assert budget.total_units == [32 * exact_units_from_bits(bits[0, 0])]
    bad = ExactBudget(1, 1, 1)
    inject(bad, np.array([[0x7f800000]], dtype=np.uint32).view(np.float32))
    with pytest.raises(FloatingPointError, match='Nonfinite'):
        bad.flush()


def test_oracle_double_double_sum_tracks_exact_sum(lib):
    """The C oracle's per-cell budget accumulation against exact rationals (seed 48193 patterns)."""
    rng = np.random.default_rng(48193)
    for trial in range(20):
        bits = rng.integers(0, 1 << 32, size=4096, dtype=np.uint32)
        exps = (bits >> 23) & 255
        bits[exps == 255] ^= np.uint32(1 << 23)
        if trial % 2:   # physically scaled volumes: exponents within a few decades
            bits = (bits & np.uint32(0x807fffff)) | (rng.integers(100, 140, size=4096).astype(np.uint32) << 23)
        values = bits.view(np.float32)
        out = np.zeros(2)
        lib.nhf_oracle_dd_sum(values.ctypes.data, len(values), out.ctypes.data)
        exact = sum(Fraction(float(v)) for v in values)
        got = Fraction(out[0]) + Fraction(out[1])
        magnitude = max(abs(Fraction(float(v))) for v in values)
        assert abs(got - exact) <= magnitude * len(values) * Fraction(1, 1 << 104)
        if trial % 2:
            assert float(got) == float(exact)


# ------------------------------------------------------------------------------ faces

@pytest.mark.parametrize('length', [0.1, 0.7, 1.3, 2.2, 3.14159])
def test_single_side_face_is_the_side_value(lib, length):
    rng = np.random.default_rng(8726)
    q = np.r_[np.float32(0), rng.normal(size=64).astype(np.float32)]
    lengths = np.full(len(q), length, dtype=np.float32)
    ids = np.arange(1, len(q), dtype=np.int32)
    out = np.zeros(1, np.float32)
    for i in range(len(ids)):
        lib.nhf_oracle_face_discharge(q.ctypes.data, lengths.ctypes.data, ids.ctypes.data, i, 1, out.ctypes.data)
        assert out.view(np.uint32)[0] == q[ids[i]:ids[i] + 1].view(np.uint32)[0]


def test_multiside_face_is_length_weighted(lib):
    q = np.array([0, 1.25, -0.5, 2.75, -3, 0.125, 5, -1, 0.75, -2], dtype=np.float32)
    lengths = np.array([1, 0.3, 1.7, 0.2, 0.8, 1.5, 0.4, 2.5, 0.6, 1.3], dtype=np.float32)
    groups = [[1, 2], [3, 4, 5], [6, 7], [8, 9], []]
    flat = np.array([s for g in groups for s in g] or [0], dtype=np.int32)
    starts = [0, 2, 5, 7, 9]
    out = np.zeros(1, np.float32)
    for face, group in enumerate(groups):
        lib.nhf_oracle_face_discharge(q.ctypes.data, lengths.ctypes.data, flat.ctypes.data, starts[face],
                                      len(group), out.ctypes.data)
        if group:
            expected = np.average(q[group].astype(np.float64), weights=lengths[group].astype(np.float64))
            assert out[0] == pytest.approx(expected, abs=2e-6)
            assert out[0] != q[group[0]]
        else:
            assert out[0] == 0


def test_boundary_stencil_continues_uniform_and_linear_discharge(lib):
    cases = np.array([[1.3, 0., .8, 1., 0.], [-2.5, -1., 0., 0., 1.],
                      [1.3, .75 * 1.3, .5 * 1.3, .25, .5], [.7, 0., 0., 1., 1.]], dtype=np.float32)
    out = np.zeros(len(cases), np.float32)
    lib.nhf_oracle_neighbor_sum(cases.ctypes.data, len(cases), out.ctypes.data)
    np.testing.assert_allclose(out, 2 * cases[:, 0], rtol=2e-7, atol=2e-7)


def test_closed_stencil_keeps_left_plus_right_bits(lib):
    rng = np.random.default_rng(9041)
    values = rng.normal(size=(64, 3)).astype(np.float32)
    cases = np.zeros((64, 5), np.float32)
    cases[:, :3] = values
    out = np.zeros(64, np.float32)
    lib.nhf_oracle_neighbor_sum(cases.ctypes.data, 64, out.ctypes.data)
    np.testing.assert_array_equal(out.view(np.uint32), (values[:, 1] + values[:, 2]).view(np.uint32))


# ------------------------------------------------------------------------------ pressure

def _split_face_mesh():
    """Coarse cell 1 with two right and two upper subedges (test_pressure_gradient.py:154-201)."""
    from nh_flood_2d_amd.hydraulics import hanging_geometry
    x = np.array([0, 4, 10, 10, 4, 2, 6], dtype=np.float32)
    y = np.array([0, 4, 2, 6, -4, 10, 10], dtype=np.float32)
    side_x = np.array([0, 8, 8, 4, 2, 6], dtype=np.float32)
    side_y = np.array([0, 2, 6, 0, 8, 8], dtype=np.float32)
    topo = np.array([[0, 0, 0], [2, 1, 2], [2, 1, 3], [1, 4, 1], [1, 1, 5], [1, 1, 6]], dtype=np.int32)
    geometry = hanging_geometry(x, y, side_x, side_y, topo)
    np.testing.assert_array_equal(geometry.cells, [1])
    np.testing.assert_array_equal(geometry.axes, [[1, 1]])
    counts = np.zeros((7, 4), np.int32)
    counts[1] = [0, 2, 1, 2]
    packed = (counts[:, 0] | counts[:, 1] << 8 | counts[:, 2] << 16 | counts[:, 3] << 24).astype(np.uint32)
    arrays = dict(
        x=x, y=y, bed=np.zeros(7, np.float32), depth=np.zeros(7, np.float32),
        csr_start=np.zeros(7, np.int32), csr_count=packed, isl=np.array([1, 2, 3, 4, 5], np.int32),
        side_cells=np.ascontiguousarray(topo[:, 1:]).ravel().astype(np.int32),
        side_meta=topo[:, 0].astype(np.uint8), length=np.array([0, 4, 4, 8, 4, 4], np.float32),
        edge_index=geometry.edge_index, hang_offset=geometry.offsets.ravel(),
        hang_slot=geometry.slots.ravel(), grad_cell=geometry.cells,
        grad_axes=(geometry.axes[:, 0] | geometry.axes[:, 1] << 1).astype(np.int32))
    m = abi.MeshStruct()
    m.e_num, m.s_num, m.r_len = 7, 6, 5
    m.n_grad, m.n_hang = len(geometry.cells), len(geometry.offsets)
    for name, a in arrays.items():
        setattr(m, name, a.ctypes.data)
    return m, arrays


def test_split_face_reconstruction_table(lib):
    m, arrays = _split_face_mesh()
    states = [
        ([1., 1., 1., .5, 1.5, .25], .5 / 8, [.125, -.125]),
        ([1., 1., 1., .5, 2.0, .25], .5 / 8, [.125, -.125]),
        ([1., 1., 1., .5, 1.5, 1.75], 1.125 / 14, [2 * 1.125 / 14, -2 * 1.125 / 14]),
        ([1., 1., 1., .25, 1.5, 1.75], .625 / 6, [2 * .625 / 6, -2 * .625 / 6]),
        ([.25, 1., 1., .5, 1.5, 1.75], 0., [.75, .75]),
        ([1.25, 1., 1., .5, 1.5, .25], .75 / 8, [-.25 + 2 * .75 / 8, -.25 - 2 * .75 / 8]),
        ([1.25, .25, 1., .5, 1.5, .25], .75 / 8, [-1., -.25 - 2 * .75 / 8]),
        ([1.25, 1., 1., .25, .25, .25], 0., [-.25, -.25]),
    ]
    sides = np.array([1, 2], np.int32)
    for depths, gradient, difference in states:
        arrays['depth'][:] = [0, *depths]
        raw = np.array([arrays['depth'][2] - arrays['depth'][1], arrays['depth'][3] - arrays['depth'][1]],
                       dtype=np.float32)
        grads = np.zeros(2, np.float32)
        diffs = np.zeros(2, np.float32)
        assert lib.nhf_oracle_pressure(ctypes.byref(m), np.float32(.25), 2, sides.ctypes.data, raw.ctypes.data,
                                       grads.ctypes.data, diffs.ctypes.data) == 0
        assert grads[1] == pytest.approx(gradient, abs=2e-7), depths
        np.testing.assert_allclose(diffs, difference, atol=3e-7, rtol=0)


def test_deterministic_pow_and_exp_round_like_the_correct_f32_result():
    """x^(7/3) (friction) and e^x (Horton) of the oracle, which the HIP kernels evaluate with the same
    FP64 operations: against the F32 rounding of the FP64 libm result on 200k samples each, every
    result within 1 ulp and at most 1 in 10^4 not the correctly rounded one."""
    lib = oracle.load('f32')
    lib.nhf_oracle_math.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p]
    rng = np.random.default_rng(731)
    n = 200_000
    x = (np.float32(10.0) ** rng.uniform(-5.0, 2.4, n)).astype(np.float32)
    x[:4] = [1e-5, 1.0, 0.0, 2.0]
    y = np.full(n, np.float32(7.0) / np.float32(3.0), dtype=np.float32)
    pow_out = np.zeros(n, np.float32)
    exp_out = np.zeros(n, np.float32)
    lib.nhf_oracle_math(x.ctypes.data, y.ctypes.data, n, pow_out.ctypes.data, exp_out.ctypes.data)
    ref = np.array([math.pow(float(a), float(y[0])) for a in x]).astype(np.float32)
    ulp = np.abs(pow_out.view(np.int32).astype(np.int64) - ref.view(np.int32).astype(np.int64))
    assert ulp.max() <= 1 and np.count_nonzero(ulp) <= n // 10_000
    assert pow_out[1] == 1.0 and pow_out[2] == 0.0
    z = (-rng.uniform(0.0, 40.0, n)).astype(np.float32)
    lib.nhf_oracle_math(x.ctypes.data, z.ctypes.data, n, pow_out.ctypes.data, exp_out.ctypes.data)
    ref = np.array([math.exp(float(a)) for a in z]).astype(np.float32)
    ulp = np.abs(exp_out.view(np.int32).astype(np.int64) - ref.view(np.int32).astype(np.int64))
    assert ulp.max() <= 1 and np.count_nonzero(ulp) <= n // 10_000
