"""Pins of the oracle's frame post-process (FLIPMethod::postProcess, base/source/FLIPMethod.cu:300-405)
by closed forms and invariants:

* the surface/velocity splat (Rasterization.cu:311-393): for one particle the weight of a cell is
  max(0, (1 - (d/1.5)^2)^3) over the 3x3x3 cells around (int)p, the level set is the distance to the
  particle minus 0.5 and the velocity the particle's; for two particles the level set measures the
  distance to their weight-averaged position;
* the redistancing (Rasterization.cu:395-490): a unit-slope plane is a fixed point of the Godunov step
  (also with the water tank's boundary values, which continue that plane); a plane of slope 2 moves
  by -0.3 S(phi) (|grad phi| - 1) exactly; a sphere's distance scaled by 3 relaxes to the distance
  itself near the interface without moving the zero level set;
* curl, normalised gradient and divergence (FieldOP.cu:80-110, :273-301, :338-369) of linear fields;
* the per-particle vorticity update (FLIPMethod.cu:372-387) inside a uniformly moving blob, where the
  curl vanishes: vort <- vort (1 - (1 - vortPresRate)^dt).
"""
import numpy as np
import pytest

from oracle import oracle as O
from test_oracle_boundary import Frame, FLUID

FIELD_SURF = O.FIELD_COUNT


def frame(n=12, **kw):
    return Frame(n=n, closed=True, fp64=True, **kw)


def cell_centres(f):
    return f.ijk + 0.5


def test_splat_single_particle_closed_form():
    f = frame()
    p = np.array([[5.3, 6.8, 4.1]], np.float32)
    v = np.array([[1.5, -2.0, 0.25]], np.float32)
    f.o.set_particles(p, v)
    f.o.post_part(0)
    surf, velo = f.o.field(FIELD_SURF), f.o.field(O.FIELD_VELO)
    c = cell_centres(f)
    d = np.linalg.norm(c - p.astype(np.float64), axis=1)
    near = np.all(np.abs(f.ijk - np.floor(p).astype(np.int64)) <= 1, axis=1)      # neighborCells, offset 1
    w = np.where(near, np.maximum(0.0, (1 - (d / 1.5) ** 2) ** 3), 0.0)
    hit = w > 1e-30
    assert 0 < hit.sum() < 27                    # the corners of the 3^3 stencil lie beyond 1.5
    np.testing.assert_allclose(surf[hit], d[hit] - 0.5, rtol=0, atol=1e-12)
    np.testing.assert_array_equal(surf[~hit], 5.0)
    np.testing.assert_allclose(velo[hit], np.broadcast_to(v.astype(np.float64), (hit.sum(), 3)), rtol=1e-12)
    np.testing.assert_array_equal(velo[~hit], 0.0)


def test_splat_two_particles_weighted_mean_position():
    f = frame()
    p = np.array([[5.3, 6.2, 4.7], [6.1, 5.9, 5.2]], np.float32)
    v = np.array([[1.0, 0.0, 0.0], [0.0, 3.0, 0.0]], np.float32)
    f.o.set_particles(p, v)
    f.o.post_part(0)
    surf, velo = f.o.field(FIELD_SURF), f.o.field(O.FIELD_VELO)
    c = cell_centres(f)
    P = p.astype(np.float64)
    ws = []
    for q in range(2):
        d = np.linalg.norm(c - P[q], axis=1)
        near = np.all(np.abs(f.ijk - np.floor(P[q]).astype(np.int64)) <= 1, axis=1)
        ws.append(np.where(near, np.maximum(0.0, (1 - (d / 1.5) ** 2) ** 3), 0.0))
    W = ws[0] + ws[1]
    both = (ws[0] > 0) & (ws[1] > 0)
    assert both.sum() > 4
    mean = (ws[0][:, None] * P[0] + ws[1][:, None] * P[1]) / np.where(W > 0, W, 1)[:, None]
    np.testing.assert_allclose(surf[both], np.linalg.norm(mean[both] - c[both], axis=1) - 0.5, rtol=0, atol=1e-12)
    vm = (ws[0][:, None] * v[0] + ws[1][:, None] * v[1]) / np.where(W > 0, W, 1)[:, None]
    np.testing.assert_allclose(velo[both], vm[both], rtol=1e-12)


@pytest.mark.parametrize("sgn", [1.0, -1.0])
def test_redistance_unit_plane_is_fixed(sgn):
    f = frame()
    c = cell_centres(f)
    phi = sgn * (c[:, 1] - 5.25)
    f.o.set_field(FIELD_SURF, phi)
    f.o.post_part(2, 0)
    np.testing.assert_array_equal(f.o.field(FIELD_SURF), phi)


def test_redistance_tank_boundary_continues_the_plane():
    # with the water tank the domain faces hold cell.y - level (and -1 / +1 below / above): the plane
    # phi = y_centre - level is again a fixed point, boundary cells included
    level = 4.75
    f = frame(enable_water_tank=1, wall_level=level / 12.0)
    c = cell_centres(f)
    phi = c[:, 1] - np.float32(level / 12.0) / np.float32(1.0 / 12.0)
    f.o.set_field(FIELD_SURF, phi)
    f.o.post_part(2, 1)
    np.testing.assert_allclose(f.o.field(FIELD_SURF), phi, rtol=0, atol=1e-12)


def test_redistance_slope_two_plane_one_step():
    f = frame()
    c = cell_centres(f)
    phi = 2.0 * (c[:, 1] - 5.25)
    f.o.set_field(FIELD_SURF, phi)
    f.o.post_part(2, 0)
    s = phi / np.sqrt(phi * phi + 1.0)
    dt = float(np.float32(0.3))                   # the reference's 0.3f
    np.testing.assert_allclose(f.o.field(FIELD_SURF), phi - dt * s * (2.0 - 1.0), rtol=0, atol=1e-12)


def test_redistance_relaxes_a_scaled_sphere_to_its_distance():
    # 80 steps (the default redistIteration) from three times a sphere's signed distance: within two
    # cells of the interface the gradient magnitude relaxes from 3 to 1 (mean within 2%), the field
    # stays within 0.75 cells of the distance (the first-order scheme lets the interface drift a
    # little) and keeps its sign away from the interface
    n = 24
    f = frame(n=n)
    c = cell_centres(f)
    centre, R = np.array([12.2, 11.6, 12.4]), 5.3
    dist = np.linalg.norm(c - centre, axis=1) - R
    f.o.set_field(FIELD_SURF, 3.0 * dist)
    for _ in range(80):
        f.o.post_part(2, 0)
    phi = f.o.field(FIELD_SURF)
    dense = np.zeros((n, n, n))
    dense[f.ijk[:, 0], f.ijk[:, 1], f.ijk[:, 2]] = phi
    grad = np.sqrt(sum(gq ** 2 for gq in np.gradient(dense)))[f.ijk[:, 0], f.ijk[:, 1], f.ijk[:, 2]]
    band = np.abs(dist) < 2.0
    assert abs(grad[band].mean() - 1.0) < 0.02 and grad[band].max() < 1.3
    assert np.abs(phi[band] - dist[band]).max() < 0.75
    far = np.abs(dist) > 1.0
    assert np.all(np.sign(phi[far]) == np.sign(dist[far]))


def test_curl_of_linear_field():
    f = frame()
    c = cell_centres(f)
    a, b, cc = 0.5, -1.25, 2.0
    v = np.stack([a * c[:, 1], b * c[:, 2], cc * c[:, 0]], 1)        # (a y, b z, c x)
    f.o.set_field(O.FIELD_VELO, v)
    f.o.post_part(3)
    np.testing.assert_array_equal(f.o.field(O.FIELD_COMP), np.broadcast_to([-b, -cc, -a], v.shape))


def test_normal_gradient_and_divergence_of_linear_fields():
    f = frame()
    c = cell_centres(f)
    g = np.array([0.5, -0.25, 1.0])
    f.o.set_field(FIELD_SURF, c @ g + 0.75)
    f.o.post_part(4)
    np.testing.assert_allclose(f.o.field(O.FIELD_COMP), np.broadcast_to(g / np.linalg.norm(g), c.shape), rtol=1e-12)
    al = np.array([0.5, -1.5, 0.25])
    f.o.set_field(O.FIELD_COMP, c * al)
    f.o.post_part(5)
    # the divergence lands in the pressure field, which post_part leaves in place
    div = f.o.field(O.FIELD_PRES)
    np.testing.assert_array_equal(div, np.full(div.shape, al.sum()))


def test_vorticity_decay_inside_uniform_blob():
    # a blob moving uniformly: the splatted velocity is uniform around interior particles, its curl 0
    # there, so their vorticity only decays: vort (1 - (1 - vortPresRate)^dt)
    n = 24
    prm = dict(ext_force=(0.0, -9.8, 0.0), wall_level=1e30, voxel_size=1.0 / n, dt=1.0 / 24.0, vort_pres_rate=0.2)
    o = O.Oracle((n, n, n), B=4, prm=O.params(**prm), fp64=True)
    pos = O.seed_box(n, n, n, 18.0, 18.0, 18.0, 8, 3)
    pos = pos[np.all(pos > 6.0, axis=1)]
    vel = np.tile(np.float32([200.0, -100.0, 50.0]), (pos.shape[0], 1))     # voxel units: a strong fringe curl
    vort = np.full(pos.shape[0], 1.0, np.float32)
    o.set_particles(pos, vel, vort=vort)
    o.post_process()
    d = o.get_particles()
    inner = np.all((pos > 9.0) & (pos < 15.0), axis=1)
    assert inner.sum() > 100
    dis = np.float32(0.8) ** np.float32(1.0 / 24.0)
    np.testing.assert_allclose(d["vort"][inner], 1.0 - float(dis), rtol=1e-6)
    # at the blob's fringe the splatted velocity falls to 0: there the curl (world units, less 0.2)
    # exceeds the decayed value
    assert (d["vort"][~inner] > 1.0 - float(dis) + 1e-3).any()
    # curvature: finite everywhere
    assert np.all(np.isfinite(d["curv"]))


@pytest.mark.parametrize("fp64", [False, True])
def test_post_process_is_splat_then_the_rest(fp64):
    # part 6 (everything after the splat) composes with the splat to the whole post-process
    n = 16
    prm = dict(ext_force=(0.0, -9.8, 0.0), wall_level=0.4, enable_water_tank=1, voxel_size=1.0 / n)
    pos = O.seed_box(n, n, n, 10.0, 9.0, 10.0, 8, 5)
    vel = np.tile(np.float32([1.0, -2.0, 0.5]), (pos.shape[0], 1))
    runs = []
    for split in (False, True):
        o = O.Oracle((n, n, n), B=4, prm=O.params(**prm), fp64=fp64)
        o.set_particles(pos, vel, vort=np.full(pos.shape[0], 0.5, np.float32))
        if split:
            o.stage(0)
            o.post_part(0)
            o.post_part(6)
        else:
            o.post_process()
        d = o.get_particles()
        runs.append([o.field(FIELD_SURF), o.field(O.FIELD_COMP), d["vort"], d["curv"]])
    for a, b in zip(*runs):
        np.testing.assert_array_equal(a, b)
