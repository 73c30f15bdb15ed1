"""Closed-form pins of the oracle's boundary rules: the right-hand side b of
FluidProjection::buildLinearSystem (base/source/FluidProjection.cu:206-244) and the pressure scatter,
air extrapolation and gradient subtract of FluidProjection::solve (:297-378).

Each test builds a small frame with every block active, writes the cell types and velocities (or the
solution x) directly, and runs one stage of the oracle.  The expected values come from mathematics,
not from the oracle's formulas:

* a linear velocity field has the same divergence at every fluid cell when all non-fluid faces are
  air, because the air-face and domain-edge rules are linear extrapolations (exact on linear fields);
* a linear pressure field has the exact gradient at every fluid cell that has a fluid neighbour
  along each axis when all other faces are solid, because the solid mirror 2p - p_opp turns the
  central difference into the one-sided one;
* a hydrostatic column (gravity-only velocity under closed walls, air above) has the linear
  pressure p_j = g dt (H - j - 1/2) and a projected velocity of zero;
* hand-derived known answers for the wall faces (+|v| / -|v|), the air-cell mean and the
  antisymmetric air ghost -p_air.

tests/test_oracle_mutants.py rebuilds the oracle with single sign/index mistakes planted in these
rules and checks that each mistake fails one of the tests here.
"""
import numpy as np
import pytest

from oracle import oracle as O

AIR, SOLID, FLUID = 0, 1, 4


class Frame:
    """An n^3 oracle frame (B=4) with every block active and direct access to its cell fields."""

    def __init__(self, n=12, closed=True, fp64=True, **kw):
        prm = dict(ext_force=(0.0, -9.8, 0.0), wall_level=1e30 if closed else -1e30, voxel_size=1.0 / n,
                   solver_mode=1, cg_rel_tol=1e-14, max_iterations=20000)
        prm.update(kw)
        self.n = n
        self.o = O.Oracle((n, n, n), B=4, prm=O.params(**prm), fp64=fp64)
        # one particle per block activates every block
        b = (np.arange(n // 4) * 4 + 2.25).astype(np.float32)
        c = np.stack(np.meshgrid(b, b, b, indexing="ij"), -1).reshape(-1, 3)
        self.o.set_particles(c, np.zeros_like(c))
        self.o.stage(0)
        self.o.stage(1)
        _, coords = self.o.block_table()
        assert coords.shape[0] == (n // 4) ** 3, "every block must be active"
        nc = coords.shape[0] * 64
        loc = np.arange(nc) % 64
        blk = coords[np.arange(nc) // 64].astype(np.int64)
        self.ijk = blk * 4 + np.stack([loc % 4, (loc // 4) % 4, loc // 16], 1)
        # storage index of each (i, j, k)
        self.index = np.empty((n, n, n), np.int64)
        self.index[self.ijk[:, 0], self.ijk[:, 1], self.ijk[:, 2]] = np.arange(nc)
        self.nc = nc

    def cell(self, i, j, k):
        return int(self.index[i, j, k])

    def set_types(self, types_ijk):
        """types_ijk[i, j, k] in {AIR, SOLID, FLUID}."""
        t = types_ijk[self.ijk[:, 0], self.ijk[:, 1], self.ijk[:, 2]].astype(np.float64)
        self.o.set_field(O.FIELD_MARK, t)
        # mass > 0 exactly where fluid (the classification's own rule)
        self.o.set_field(O.FIELD_MASS, (t == FLUID).astype(np.float64))

    def set_comp(self, fn):
        """fn(x, y, z) -> (vx, vy, vz) at cell centres (arrays)."""
        c = self.ijk + 0.5
        v = np.stack(fn(c[:, 0], c[:, 1], c[:, 2]), 1).astype(np.float64)
        self.o.set_field(O.FIELD_COMP, np.ascontiguousarray(v))
        return v

    def build(self):
        self.o.stage(2)
        fluid = self.o.field(O.FIELD_FLUID).astype(np.int64)       # storage index per rank
        return fluid, self.o.field(O.FIELD_B), self.o.field(O.FIELD_AVAL).reshape(-1, 7)

    def comp(self):
        return self.o.field(O.FIELD_COMP)

    def pres(self):
        return self.o.field(O.FIELD_PRES)


def _random_types(n, rng, p_fluid):
    return np.where(rng.random((n, n, n)) < p_fluid, FLUID, AIR)


# ------------------------------------------------------------------ A8: right-hand side b
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_b_linear_field_open_faces_is_minus_divergence(seed):
    """FluidProjection.cu:220-235 with every non-fluid face air (open domain, wallLevel=-1e30):
    the air ghost 2v - v_opp and the domain-edge value 2v - v_inner are linear extrapolations, so
    for v = (a x + d, b y + e, c z + f) every fluid cell gets b = -(a + b + c) exactly, whatever
    its neighbours.  Dyadic coefficients keep every intermediate exact in FP64."""
    rng = np.random.default_rng(seed)
    f = Frame(closed=False)
    types = _random_types(f.n, rng, 0.55)
    f.set_types(types)
    a, b, c = 0.75, -1.25, 0.625
    f.set_comp(lambda x, y, z: (a * x + 0.125, b * y - 2.0, c * z + 3.5))
    fluid, rhs, _ = f.build()
    assert fluid.size == int((types == FLUID).sum())
    ijk = f.ijk[fluid]
    # the pattern must exercise every rule: air faces on both sides, and the domain edges
    assert (ijk == 0).any() and (ijk == f.n - 1).any()
    np.testing.assert_array_equal(rhs, np.full(fluid.size, -(a + b + c)))


def test_b_wall_faces_known_answer():
    """Closed walls (wallLevel above the domain): a domain face is kWall, whose ghost velocity is
    +|v| on the low side and -|v| on the high side (FluidProjection.cu:227-232).  One fluid cell in
    the low corner and one in the high corner, air elsewhere; the values are hand-derived."""
    f = Frame(closed=True)
    n = f.n
    types = np.full((n, n, n), AIR)
    types[0, 0, 0] = FLUID
    types[n - 1, n - 1, n - 1] = FLUID
    f.set_types(types)
    v = np.zeros((f.nc, 3))
    # low corner cell and its +x/+y/+z air neighbours
    v[f.cell(0, 0, 0)] = (-3.0, 5.0, -7.0)
    v[f.cell(1, 0, 0), 0] = 11.0
    v[f.cell(0, 1, 0), 1] = 13.0
    v[f.cell(0, 0, 1), 2] = 17.0
    # high corner cell and its -x/-y/-z air neighbours
    e = n - 1
    v[f.cell(e, e, e)] = (2.0, -6.0, 9.0)
    v[f.cell(e - 1, e, e), 0] = 19.0
    v[f.cell(e, e - 1, e), 1] = -23.0
    v[f.cell(e, e, e - 1), 2] = 29.0
    f.o.set_field(O.FIELD_COMP, v)
    fluid, rhs, aval = f.build()
    got = dict(zip(map(int, fluid), rhs))
    # low corner: wall face +|v|; air face 2v - (2v - v_nbr) = v_nbr (the domain-edge value of the
    # opposite side enters the air ghost):  x: (11 - 3)/2, y: (13 - 5)/2, z: (17 - 7)/2
    assert got[f.cell(0, 0, 0)] == -(4.0 + 4.0 + 5.0)
    # high corner: wall face -|v|; air face v_nbr:  x: (-2 - 19)/2, y: (-6 + 23)/2, z: (-9 - 29)/2
    assert got[f.cell(e, e, e)] == -(-10.5 + 8.5 - 19.0)
    # diagonal 6 + nAir - nWall: three air and three wall faces each
    np.testing.assert_array_equal(aval[:, 0], [6.0, 6.0])
    np.testing.assert_array_equal(aval[:, 1:], 0.0)


def test_b_solid_neighbour_uses_its_velocity_and_diagonal():
    """A solid (collision) neighbour is neither air nor wall: its stored velocity enters the central
    difference as is, and it lowers the diagonal like a wall (FluidProjection.cu:230-244)."""
    f = Frame(closed=False)
    n = f.n
    types = np.full((n, n, n), FLUID)
    types[4, 6, 6] = SOLID            # -x of (5,6,6)
    types[5, 7, 6] = AIR              # +y of (5,6,6)
    f.set_types(types)
    v = np.zeros((f.nc, 3))
    v[f.cell(5, 6, 6)] = (1.0, 2.0, 3.0)
    v[f.cell(4, 6, 6), 0] = -8.0      # the solid cell's velocity
    v[f.cell(6, 6, 6), 0] = 4.0
    v[f.cell(5, 5, 6), 1] = -1.0
    v[f.cell(5, 6, 5), 2] = 6.0
    v[f.cell(5, 6, 7), 2] = 10.0
    f.o.set_field(O.FIELD_COMP, v)
    fluid, rhs, aval = f.build()
    r = int(np.nonzero(fluid == f.cell(5, 6, 6))[0][0])
    # x: (4 - (-8))/2 = 6;  y (air above): (2*2 - (-1)) - (-1) = 6 -> 3;  z: (10 - 6)/2 = 2
    assert rhs[r] == -(6.0 + 3.0 + 2.0)
    assert aval[r, 0] == 6.0 + 1.0 - 1.0          # one air (+1), one solid (-1)
    # off-diagonals: -1 for the 4 fluid neighbours (+x, -y, -z, +z), 0 for -x (solid) and +y (air)
    np.testing.assert_array_equal(aval[r, 1:], [0.0, -1.0, -1.0, 0.0, -1.0, -1.0])


# ------------------------------------------------------------------ A11: pressure apply
def _set_x_from_cells(f, fluid, p_of_cell):
    x = np.array([p_of_cell[int(c)] for c in fluid], np.float64)
    f.o.set_field(O.FIELD_X, x)


@pytest.mark.parametrize("seed", [0, 1])
def test_gradient_of_linear_pressure_with_solid_mirrors(seed):
    """FluidProjection.cu:349-375 with closed walls and solid cells but no air: the mirror ghost
    2p - p_opp at a solid or wall face makes the central difference one-sided, so a linear pressure
    p = a i + b j + c k + d gives the exact gradient (a, b, c) at every fluid cell that has a fluid
    neighbour along each axis.  Solid cells sit on a lattice of spacing 3 (including index 0), so
    no fluid cell is enclosed on both sides along an axis."""
    rng = np.random.default_rng(seed)
    f = Frame(closed=True)
    n = f.n
    types = np.full((n, n, n), FLUID)
    lat = np.array([0, 3, 6, 9])
    sel = rng.random((lat.size,) * 3) < 0.6
    for (ii, jj, kk) in zip(*np.nonzero(sel)):
        types[lat[ii], lat[jj], lat[kk]] = SOLID
    f.set_types(types)
    v0 = f.set_comp(lambda x, y, z: (np.sin(x), np.cos(y), x * 0 + 0.25))
    fluid, _, _ = f.build()
    a, b, c = 0.5, -1.5, 2.25
    ijk = f.ijk
    p_cell = a * ijk[:, 0] + b * ijk[:, 1] + c * ijk[:, 2] + 7.0
    _set_x_from_cells(f, fluid, p_cell)
    f.o.apply_pressure()
    comp = f.comp()
    fl = fluid
    np.testing.assert_array_equal(comp[fl], v0[fl] - np.array([a, b, c]))
    # solid cells keep their velocity and get no pressure
    solid = np.nonzero(types[ijk[:, 0], ijk[:, 1], ijk[:, 2]] == SOLID)[0]
    assert solid.size > 5
    np.testing.assert_array_equal(comp[solid], v0[solid])
    np.testing.assert_array_equal(f.pres()[solid], 0.0)


def test_air_mean_and_air_ghost_known_answer():
    """K11 (FluidProjection.cu:314-335): an air cell takes the mean pressure of its fluid face
    neighbours (unchanged 0 without any); K12 (:360-375): across an air face the ghost pressure is
    -p_air, across a solid face 2p - p_opp.  Hand-derived values for three fluid cells."""
    f = Frame(closed=True)
    n = f.n
    types = np.full((n, n, n), AIR)
    F1, F2, F3 = (4, 4, 4), (6, 4, 4), (5, 5, 4)
    for c in (F1, F2, F3):
        types[c] = FLUID
    types[3, 4, 4] = SOLID            # -x of F1
    f.set_types(types)
    v = np.zeros((f.nc, 3))
    f.o.set_field(O.FIELD_COMP, v)
    fluid, _, _ = f.build()
    p_cell = np.zeros(f.nc)
    p_cell[f.cell(*F1)], p_cell[f.cell(*F2)], p_cell[f.cell(*F3)] = 2.0, 6.0, 10.0
    _set_x_from_cells(f, fluid, p_cell)
    f.o.apply_pressure()
    P = f.pres()
    # air means: (5,4,4) sees F1, F2, F3 -> 6; (4,5,4) sees F1 (-y) and F3 (+x) -> 6;
    # (4,4,5) sees F1 only -> 2; (6,5,4) sees F2 (-y) and F3 (-x) -> 8; far air keeps 0
    assert P[f.cell(5, 4, 4)] == 6.0
    assert P[f.cell(4, 5, 4)] == 6.0
    assert P[f.cell(4, 4, 5)] == 2.0
    assert P[f.cell(6, 5, 4)] == 8.0
    assert P[f.cell(9, 9, 9)] == 0.0
    assert P[f.cell(3, 4, 4)] == 0.0              # the solid cell is not extrapolated
    comp = f.comp()
    # F1: x: solid at -x -> 2p - p(+x air mean 6) = -2; air at +x -> -6; (-6 - (-2))/2 = -2
    #     y: air -y (mean 2) -> -2; air +y (mean 6) -> -6: (-6 + 2)/2 = -2
    #     z: air -z (mean 2) -> -2; air +z (mean 2) -> -2: 0
    np.testing.assert_array_equal(comp[f.cell(*F1)], [2.0, 2.0, 0.0])
    # F2 = (6,4,4): x: air -x (mean 6) -> -6; air +x (7,4,4) (F2 only -> 6) -> -6: 0
    #     y: air -y (6,3,4) (F2 only -> 6) -> -6; air +y (6,5,4) (mean 8) -> -8: (-8 + 6)/2 = -1
    #     z: both air with F2 only -> -6, -6: 0
    np.testing.assert_array_equal(comp[f.cell(*F2)], [0.0, 1.0, 0.0])
    # F3 = (5,5,4): x: air -x (4,5,4) mean 6 -> -6; air +x (6,5,4) mean 8 -> -8: -1
    #     y: air -y (5,4,4) mean 6 -> -6; air +y (5,6,4) (F3 only -> 10) -> -10: -2
    #     z: air with F3 only -> -10, -10: 0
    np.testing.assert_array_equal(comp[f.cell(*F3)], [1.0, 2.0, 0.0])


def test_domain_edge_pressure_extrapolation_known_answer():
    """At a domain face the neighbour value is extrapolated, 2p - p_inner (FluidProjection.cu:
    360-365), and then takes the face's rule: with open faces (air) the ghost is -(2p - p_inner).
    A fluid cell at i = 0 with a fluid neighbour at i = 1: grad.x = (p1 - (-(2 p0 - p1)))/2 = p0."""
    f = Frame(closed=False)
    n = f.n
    types = np.full((n, n, n), AIR)
    types[0, 5, 5] = FLUID
    types[1, 5, 5] = FLUID
    f.set_types(types)
    f.o.set_field(O.FIELD_COMP, np.zeros((f.nc, 3)))
    fluid, _, _ = f.build()
    p_cell = np.zeros(f.nc)
    p_cell[f.cell(0, 5, 5)], p_cell[f.cell(1, 5, 5)] = 3.0, 7.0
    _set_x_from_cells(f, fluid, p_cell)
    f.o.apply_pressure()
    comp = f.comp()
    # (0,5,5): x: -(2*3 - 7) = 1 at -x; +x fluid 7: (7 - 1)/2 = 3
    #          y, z: air cells with only this fluid neighbour -> mean 3 -> ghost -3 on both sides: 0
    np.testing.assert_array_equal(comp[f.cell(0, 5, 5)], [-3.0, 0.0, 0.0])
    # (1,5,5): x: -x fluid 3; +x air (2,5,5) mean 7 -> -7: (-7 - 3)/2 = -5
    np.testing.assert_array_equal(comp[f.cell(1, 5, 5)], [5.0, 0.0, 0.0])
    # closed walls: the domain face is solid, mirror 2p - p_inner: (7 - (2*3 - 7))/2 = 4 = p1 - p0
    g = Frame(closed=True)
    g.set_types(types)
    g.o.set_field(O.FIELD_COMP, np.zeros((g.nc, 3)))
    fluid, _, _ = g.build()
    _set_x_from_cells(g, fluid, p_cell)
    g.o.apply_pressure()
    np.testing.assert_array_equal(g.comp()[g.cell(0, 5, 5)], [-4.0, 0.0, 0.0])


# ------------------------------------------------------------------ A8 + A9 + A11: hydrostatics
@pytest.mark.parametrize("H", [3, 7])
def test_hydrostatic_column(H):
    """Closed walls, fluid in every cell with j < H, air above, velocity (0, -g dt, 0) (gravity
    only).  Then b = g dt on the bottom row and 0 elsewhere, the pressure is hydrostatic,
    p_j = g dt (H - j - 1/2) (zero at the free surface, the face between the top fluid row and the
    air), and the projection cancels the velocity at every fluid cell."""
    f = Frame(closed=True)
    n = f.n
    types = np.full((n, n, n), AIR)
    types[:, :H, :] = FLUID
    f.set_types(types)
    gdt = 9.8 / 240.0
    f.set_comp(lambda x, y, z: (0 * x, 0 * y - gdt, 0 * z))
    fluid, rhs, _ = f.build()
    j = f.ijk[fluid, 1]
    np.testing.assert_allclose(rhs, np.where(j == 0, gdt, 0.0), rtol=0, atol=1e-15)
    f.o.stage(3)
    P = f.pres()
    np.testing.assert_allclose(P[fluid], gdt * (H - j - 0.5), rtol=1e-9)
    comp = f.comp()
    np.testing.assert_allclose(comp[fluid], 0.0, rtol=0, atol=1e-9 * gdt)
