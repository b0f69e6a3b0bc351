"""Meshes of different sizes concatenated into one (union.py), checked on the CPU.

The concatenation is itself a mesh: stepped as one mesh in the oracle, with a dt forced below
every part's CFL limit (so one clock takes the step every part would take alone), each real cell
and side must give the same bits as its part stepped alone. That pins every index offset of the
builder (side cells, side lists, cell records, boundary ids, rain pixels, land classes,
hanging-edge data). The device's per-part clocks are the GPU test's subject.
"""
import numpy as np
import pytest

import oracle
from nh_flood_2d_amd import abi, synth
from nh_flood_2d_amd.clock import seconds_to_ticks
from nh_flood_2d_amd.engine import Engine, make_params
from nh_flood_2d_amd.forcing import cell_forcing_arrays
from nh_flood_2d_amd.layout import Layout, build_layout, initial_state
from nh_flood_2d_amd.union import UnionPlan
from parity import coefficients_at


def parts(mixed: bool):
    scs = [synth.small_multires(1, blocks=20, pond=0.3), synth.small_multires(2, blocks=12, pond=0.3, discharge=0.5)]
    if mixed:
        scs.insert(1, synth.plain_uniform(3, n=24, dx=10.))
    out = []
    for sc in scs:
        forcing = synth.scenario_forcing(sc)
        layout = build_layout(sc.mesh, sc.land_parameters, cell_forcing_arrays(sc.mesh, forcing), sc.sita)
        state = initial_state(sc.mesh, layout)
        if sc.initial_depth is not None:
            state['depth'] = sc.initial_depth.astype(np.float32).copy()
        out.append((sc, forcing, layout, state))
    return out


def union_layout(plan: UnionPlan) -> Layout:
    """The concatenation as a Layout of its own (identity numbering)."""
    cells, sides = np.arange(plan.e_num), np.arange(plan.s_num)
    arrays = dict(plan.arrays)
    arrays['_r_len'] = np.array([plan.r_len])
    return Layout(plan.e_num, plan.s_num, cells, cells, sides, sides, arrays, plan.n_pixels, plan.b_num,
                  plan.n_grad, plan.n_hang, plan.theta, size_base=plan.size_base)


def union_state(plan: UnionPlan, items) -> dict:
    out = {}
    for name, values in items[0][3].items():
        n = plan.s_num if name == 'q' else plan.e_num
        arr = np.zeros(n, dtype=values.dtype)
        for p, (_, _, layout, state) in enumerate(items):
            if name == 'q':
                arr[plan.side_begin[p]:plan.side_begin[p] + layout.s_num - 1] = layout.sides_to_device(state[name])[1:]
            else:
                arr[plan.cell_begin[p]:plan.cell_begin[p] + layout.e_num - 1] = layout.cells_to_device(state[name])[1:]
        out[name] = arr
    return out


def test_plan_ranges_and_offsets():
    items = parts(mixed=True)
    plan = UnionPlan([it[2] for it in items])
    assert plan.e_num - 1 == sum(-(-(lay.e_num - 1) // 1024) * 1024 for *_, lay, _ in items)
    assert plan.size_base == 0.0            # 2 m and 10 m bases are not powers of two apart
    table = plan.part_struct()
    for p, (_, _, layout, _) in enumerate(items):
        t = table[p]
        assert (t.cell_begin - 1) % 1024 == 0 and (t.side_begin - 1) % 1024 == 0
        assert t.cell_end - t.cell_begin == layout.e_num - 1 and t.side_end - t.side_begin == layout.s_num - 1
        sc = plan.arrays['side_cells'].reshape(-1, 2)[t.side_begin:t.side_end]
        local = layout.arrays['side_cells'].reshape(-1, 2)[1:]
        back = np.where(sc == 0, 0, sc - t.cell_begin + 1)
        np.testing.assert_array_equal(back, local)
        assert np.all((sc == 0) | ((sc >= t.cell_begin) & (sc < t.cell_end)))
    same = UnionPlan([it[2] for it in items if it[2].size_base == items[0][2].size_base])
    assert same.size_base == items[0][2].size_base
    with pytest.raises(ValueError, match='sita'):
        other = parts(mixed=False)[0][2]
        other.theta = 0.8
        UnionPlan([items[0][2], other])


@pytest.mark.parametrize('mixed', [False, True])
def test_concatenation_steps_like_each_part_alone(mixed):
    items = parts(mixed)
    plan = UnionPlan([it[2] for it in items])
    sc0 = items[0][0]
    params = make_params(sc0.afa, sc0.sita, sc0.min_h, 32)
    alone = [Engine(layout, params, state, device=oracle.HostDevice('f32')) for _, _, layout, state in items]
    whole = Engine(union_layout(plan), params, union_state(plan, items), device=oracle.HostDevice('f32'))
    engines = alone + [whole]
    for e in engines:
        e.reset(0)
        e.compute_cfl()
    now = 0
    for _ in range(8):
        t = now / 2 ** 40
        coeffs, rates = [], []
        for (sc, forcing, layout, _), e in zip(items, alone):
            c = coefficients_at(forcing, t)[:layout.b_num].ravel()
            r = forcing.rain.rates(forcing.rain.index_at(t)).astype(np.float32)
            e.set_forcing(c if layout.b_num else np.zeros(0, np.float32), r, True)
            coeffs.append(c)
            rates.append(r)
        whole.set_forcing(np.concatenate(coeffs).astype(np.float32), np.concatenate(rates), True)
        step = min(0.5, 0.5 * min(e.cfl_state()[0] for e in engines))
        target = seconds_to_ticks(t + step)
        for e in engines:
            e.begin(target, now)
            rows = e.run_batch(1)
            assert rows[0, 3] == abi.COMPLETED
        now = (int(rows[0, 1]) << 32) | int(rows[0, 2])
    for p, ((sc, _, layout, _), e) in enumerate(zip(items, alone)):
        c0, s0 = int(plan.cell_begin[p]), int(plan.side_begin[p])
        for name in ('depth', 'max_depth', 'rain_time', 'capacity', 'max_speed'):
            mine = layout.cells_to_device(e.read_state(name))[1:]
            theirs = whole.read_state(name)[c0:c0 + layout.e_num - 1]
            assert mine.tobytes() == theirs.tobytes(), (p, name)
        mine = layout.sides_to_device(e.read_state('q'))[1:]
        assert mine.tobytes() == whole.read_state('q')[s0:s0 + layout.s_num - 1].tobytes(), (p, 'q')
        assert sc.initial_depth is None or np.any(mine != 0)   # the ponded parts move water
    for e in engines:
        e.close()


def test_uneven_meshes_have_the_drawn_levels_and_classes():
    """The uneven-size batch's meshes: 1-3 levels with 5 m finest cells (one size base, so a union
    keeps the sizes-from-levels path), 2-4 land classes, and about the requested cell count."""
    for levels in (1, 2, 3):
        sc = synth.uneven_scenario(levels, 6000, levels, classes=levels + 1)
        sizes = np.unique(sc.mesh.geometry()[1][1:])
        np.testing.assert_allclose(sizes, 5.0 * 2.0 ** np.arange(len(sizes)))
        assert len(sizes) == levels
        assert len(np.unique(sc.mesh.land[1:])) == levels + 1
        assert 0.6 * 6000 < sc.mesh.e_num - 1 < 1.6 * 6000
    items = synth.uneven_batch(3, 4, 2000, 9000, workers=1)
    plan = UnionPlan([lay for _, lay in items])
    assert plan.size_base == 5.0
    assert [lay.e_num for _, lay in items] == [sc.mesh.e_num for sc, _ in items]


def test_part_tables_are_checked_by_the_library():
    """nhu_workspace_bytes plans without touching a GPU: it accepts parts that tile the mesh in
    order from 1 + 1024 k and refuses gaps, overlaps, misaligned starts and wrong totals."""
    import ctypes
    lib = abi.load_hip_library()
    items = parts(mixed=False)
    plan = UnionPlan([it[2] for it in items])
    m = abi.MeshStruct()
    m.e_num, m.s_num, m.r_len = plan.e_num, plan.s_num, plan.r_len
    m.b_num, m.n_pixels, m.n_grad, m.n_hang = plan.b_num, plan.n_pixels, plan.n_grad, plan.n_hang
    params = make_params(0.5, 1.0, 0.02, 32)
    good = plan.part_struct()
    assert lib.nhu_workspace_bytes(ctypes.byref(m), ctypes.byref(params), good, len(good)) > 0

    def refused(edit):
        bad = plan.part_struct()
        edit(bad)
        return lib.nhu_workspace_bytes(ctypes.byref(m), ctypes.byref(params), bad, len(bad)) == 0
    assert refused(lambda t: setattr(t[1], 'cell_begin', t[1].cell_begin + 1))       # misaligned
    assert refused(lambda t: setattr(t[0], 'cell_end', t[0].cell_end + 2000))        # overlaps the next
    assert refused(lambda t: setattr(t[1], 'side_begin', t[1].side_begin + 1024))    # gap in the sides
    assert refused(lambda t: setattr(t[1], 'b_base', t[1].b_base + 1))              # boundaries not tiled
    assert refused(lambda t: setattr(t[0], 'pix_count', 0))                          # no rain pixel
    assert lib.nhu_workspace_bytes(ctypes.byref(m), ctypes.byref(params), good, 1) == 0   # totals short
    assert lib.nhu_workspace_bytes(ctypes.byref(m), ctypes.byref(params), good, 0) == 0


def test_concatenation_with_theta_below_one_steps_like_each_part_alone():
    """sita = 0.7 adds the face-average pass and the open-boundary fractions (per cell, 4 each)."""
    scs = [synth.small_multires(4, blocks=16, pond=0.3, theta=0.7),
           synth.small_multires(5, blocks=10, pond=0.3, theta=0.7, discharge=0.4)]
    items = []
    for sc in scs:
        forcing = synth.scenario_forcing(sc)
        layout = build_layout(sc.mesh, sc.land_parameters, cell_forcing_arrays(sc.mesh, forcing), sc.sita)
        state = initial_state(sc.mesh, layout)
        state['depth'] = sc.initial_depth.astype(np.float32).copy()
        items.append((sc, forcing, layout, state))
    plan = UnionPlan([it[2] for it in items])
    assert 'open_face' in plan.arrays
    params = make_params(0.5, 0.7, 0.02, 32)
    alone = [Engine(layout, params, state, device=oracle.HostDevice('f32')) for _, _, layout, state in items]
    whole = Engine(union_layout(plan), params, union_state(plan, items), device=oracle.HostDevice('f32'))
    engines = alone + [whole]
    for e in engines:
        e.reset(0)
        e.compute_cfl()
    now = 0
    for _ in range(6):
        t = now / 2 ** 40
        coeffs, rates = [], []
        for (sc, forcing, layout, _), e in zip(items, alone):
            c = coefficients_at(forcing, t)[:layout.b_num].ravel()
            r = forcing.rain.rates(forcing.rain.index_at(t)).astype(np.float32)
            e.set_forcing(c, r, True)
            coeffs.append(c)
            rates.append(r)
        whole.set_forcing(np.concatenate(coeffs).astype(np.float32), np.concatenate(rates), True)
        step = min(0.5, 0.5 * min(e.cfl_state()[0] for e in engines))
        target = seconds_to_ticks(t + step)
        for e in engines:
            e.begin(target, now)
            rows = e.run_batch(1)
            assert rows[0, 3] == abi.COMPLETED
        now = (int(rows[0, 1]) << 32) | int(rows[0, 2])
    for p, ((_, _, layout, _), e) in enumerate(zip(items, alone)):
        c0, s0 = int(plan.cell_begin[p]), int(plan.side_begin[p])
        assert (layout.cells_to_device(e.read_state('depth'))[1:].tobytes()
                == whole.read_state('depth')[c0:c0 + layout.e_num - 1].tobytes())
        assert (layout.sides_to_device(e.read_state('q'))[1:].tobytes()
                == whole.read_state('q')[s0:s0 + layout.s_num - 1].tobytes())
    for e in engines:
        e.close()
