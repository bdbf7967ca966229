"""Pins of the oracle functions that the reference's fixtures and golden
vectors of ordinary states leave open:

* the CFL max's NaN semantics (Heterogeneous_MHD/C/calcfl.c:55, max_2num
  of :93-103: a NaN speed makes c NaN, the next speed replaces it) and its
  non-positive densities: against the reference build and committed vectors,
  and by closed forms in which a NaN-skipping or order-blind max fails;
* the cell-averaged B^2 mode of the CFL (the averages of calcfl.c:38-43):
  closed forms on non-uniform states, one per axis on a non-cubic box, where
  averaging with the i-1 (j-1, k-1) face or along another axis gives a
  different answer;
* the FP64 monitors: exact sums of distinct per-component values (a swapped
  component fails) and a divergence whose value depends on every +1
  neighbour of the left-face stencil.
"""
import numpy as np
import pytest

import oracle_lib
from conftest import same
from golden.special_states import CFL_CASES, EXTRA_CASES, special_state

F32 = np.float32
G = 5.0 / 3.0

needs_ref = pytest.mark.skipif(not oracle_lib.reference_available(),
                               reason="oracle/_ref not built (reference tree absent)")


def f32_equal(a, b):
    a, b = F32(a), F32(b)
    return a.tobytes() == b.tobytes() or (np.isnan(a) and np.isnan(b))


# ------------------------------------------------- against the reference ----

@needs_ref
@pytest.mark.parametrize("name", CFL_CASES)
@pytest.mark.parametrize("shape", [(8, 8, 8), (12, 6, 10)])
def test_special_states_calcfl_and_ops_equal_reference(oracle, name, shape):
    ref = oracle_lib.Reference()
    u, b = special_state(name, *shape, seed=5)
    assert f32_equal(oracle.calcfl(u, b, *shape), ref.calcfl(u, b, *shape))
    for op in ("fluidx", "advectbyzx"):
        ua, ba, ub, bb = u.copy(), b.copy(), u.copy(), b.copy()
        getattr(oracle, op)(ua, ba, *shape, F32(0.2))
        getattr(ref, op)(ub, bb, *shape, F32(0.2))
        assert same(ua, ub) and same(ba, bb), op


@needs_ref
@pytest.mark.parametrize("name", CFL_CASES)
def test_special_states_steps_equal_reference(oracle, name):
    # cubes: the reference's transpose13 handles nx == nz only
    ref = oracle_lib.Reference()
    for n in (8, 10):
        u, b = special_state(name, n, n, n, seed=5)
        ua, ba, ub, bb = u.copy(), b.copy(), u.copy(), b.copy()
        da = oracle.run(ua, ba, n, n, n, 3)
        db = ref.run(ub, bb, n, n, n, 3)
        assert same(da, db) and same(ua, ub) and same(ba, bb)


@needs_ref
@pytest.mark.parametrize("name", EXTRA_CASES)
def test_extra_states_equal_reference(oracle, name):
    # states beyond the committed vectors (FP32 overflow inside the sweeps):
    # the CFL, single ops on boxes of several shapes, and three whole steps
    ref = oracle_lib.Reference()
    for shape in ((8, 8, 8), (12, 6, 10), (20, 4, 6)):
        u, b = special_state(name, *shape, seed=5)
        assert f32_equal(oracle.calcfl(u, b, *shape), ref.calcfl(u, b, *shape))
        for op in ("fluidx", "advectbyzx"):
            ua, ba, ub, bb = u.copy(), b.copy(), u.copy(), b.copy()
            getattr(oracle, op)(ua, ba, *shape, F32(0.2))
            getattr(ref, op)(ub, bb, *shape, F32(0.2))
            assert same(ua, ub) and same(ba, bb), (op, shape)
    for n in (8, 10):
        u, b = special_state(name, n, n, n, seed=5)
        ua, ba, ub, bb = u.copy(), b.copy(), u.copy(), b.copy()
        assert same(oracle.run(ua, ba, n, n, n, 3), ref.run(ub, bb, n, n, n, 3))
        assert same(ua, ub) and same(ba, bb)


def test_special_states_golden(oracle, golden):
    # vectors written by the reference build (tests/golden/make_golden.py)
    for name in CFL_CASES:
        u, b = special_state(name, 12, 6, 10, seed=5)
        assert f32_equal(oracle.calcfl(u, b, 12, 6, 10), golden[f"spc_{name}_calcfl"]), name
    for name in ("neg_density", "neg_density_many", "tiny_density", "huge_momentum"):
        u, b = special_state(name, 10, 10, 10, seed=5)
        d = oracle.run(u, b, 10, 10, 10, 3)
        assert same(d, golden[f"spc_{name}_run3_dt"]), name
        assert same(u, golden[f"spc_{name}_run3_u"]) and same(b, golden[f"spc_{name}_run3_b"]), name


# ------------------------------------------------------------ closed forms ----

def uniform(shape, rho=1.0, E=2.0, B=(0.5, 0.25, -0.375)):
    nc = int(np.prod(shape))
    u = np.tile(np.array([rho, 0, 0, 0, E], F32), nc)
    b = np.tile(np.array(B, F32), nc)
    return u, b


def speed(rho, B2, E=2.0):
    # calcfl.c:45-54 at v = 0: p = (gamma-1)(E - B2/2), c = sqrt(|2 B2 + gamma p| / rho)
    p = (G - 1.0) * (E - B2 / 2.0)
    return np.sqrt(abs(2.0 * B2 + G * p) / rho)


def test_cfl_nan_resets_the_running_max(oracle):
    shape = (6, 5, 4)
    B2 = 0.5 ** 2 + 0.25 ** 2 + 0.375 ** 2
    c0, cf = speed(1.0, B2), speed(0.25, B2)

    def case(fast, nan):
        u, b = uniform(shape)
        U = u.reshape(-1, 5)
        for s in fast:
            U[s, 0] = F32(0.25)
        for s in nan:
            U[s, 4] = F32(np.nan)
        return oracle.calcfl(u, b, *shape)

    # a fast cell before the last NaN one does not count (a NaN-skipping max
    # would give 1/cf)
    assert np.isclose(case([10], [50]), 1 / c0, rtol=1e-6)
    assert np.isclose(case([10], [50, 70]), 1 / c0, rtol=1e-6)
    # after it, it does
    assert np.isclose(case([50], [10]), 1 / cf, rtol=1e-6)
    assert np.isclose(case([10, 100], [50]), 1 / cf, rtol=1e-6)
    # no NaN: the plain max; NaN in the last cell: NaN
    assert np.isclose(case([10], []), 1 / cf, rtol=1e-6)
    assert np.isnan(case([10], [119]))
    # a negative density gives a NaN speed (sqrt of |2B^2 + gamma p| / rho < 0)
    u, b = uniform(shape)
    U = u.reshape(-1, 5)
    U[10, 0] = F32(0.25)
    U[50, 0] = F32(-1.0)
    assert np.isclose(oracle.calcfl(u, b, *shape), 1 / c0, rtol=1e-6)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_cfl_cell_averaged_field_per_axis(oracle, axis):
    # one slow-density (fast) cell whose right-hand face along `axis` carries a
    # large field: only the correct (b(i) + b(i+1))/2 average on that axis
    # brings that face into the fast cell's B^2
    nx, ny, nz = 8, 6, 5
    i0, j0, k0 = 3, 2, 1
    B0 = np.array([0.5, 0.25, -0.375])
    big = 3.0
    u, b = uniform((nx, ny, nz), B=B0)
    U, Bf = u.reshape(nz, ny, nx, 5), b.reshape(nz, ny, nx, 3)
    U[k0, j0, i0, 0] = F32(0.01)
    nb = [(k0, j0, i0 + 1), (k0, j0 + 1, i0), (k0 + 1, j0, i0)][axis]
    Bf[nb + (axis,)] = F32(big)
    rho_s = float(F32(0.01))
    face = speed(rho_s, float(B0 @ B0))
    Bavg = B0.copy()
    Bavg[axis] = (B0[axis] + big) / 2.0
    avg = speed(rho_s, float(Bavg @ Bavg))
    assert np.isclose(oracle.calcfl(u, b, nx, ny, nz, avg_b=False), 1 / face, rtol=1e-6)
    assert np.isclose(oracle.calcfl(u, b, nx, ny, nz, avg_b=True), 1 / avg, rtol=1e-6)
    assert abs(avg / face - 1) > 0.2  # the two modes are told apart


def test_monitors_closed_form(oracle):
    nx, ny, nz = 8, 6, 4
    nc = nx * ny * nz
    uc = np.array([1.25, 0.5, -0.375, 0.75, 3.5], F32)
    u = np.tile(uc, nc)
    b = np.tile(np.array([0.5, 0.25, -0.125], F32), nc)
    m = oracle.monitors(u, b, nx, ny, nz)
    # mass, energy, momentum x, y, z: exact sums of distinct values
    assert list(m[:5]) == [nc * 1.25, nc * 3.5, nc * 0.5, nc * -0.375, nc * 0.75]
    assert m[5] == 0.0
    # raise the three right-hand faces of one cell by d: the left-face
    # divergence of that cell is 3d (its +1 neighbours in x, y, z), of its
    # neighbours -d; a missing term or a wrong neighbour gives 2d or d
    d = 0.0625
    Bf = b.reshape(nz, ny, nx, 3)
    i0, j0, k0 = 2, 3, 1
    Bf[k0, j0, i0 + 1, 0] += F32(d)
    Bf[k0, j0 + 1, i0, 1] += F32(d)
    Bf[k0 + 1, j0, i0, 2] += F32(d)
    m = oracle.monitors(u, b, nx, ny, nz)
    assert m[5] == 3 * d
    # periodic wrap: the same at the box's last cell, whose +1 faces are the first ones
    b = np.tile(np.array([0.5, 0.25, -0.125], F32), nc)
    Bf = b.reshape(nz, ny, nx, 3)
    Bf[nz - 1, ny - 1, 0, 0] += F32(d)
    Bf[nz - 1, 0, nx - 1, 1] += F32(d)
    Bf[0, ny - 1, nx - 1, 2] += F32(d)
    assert oracle.monitors(u, b, nx, ny, nz)[5] == 3 * d
