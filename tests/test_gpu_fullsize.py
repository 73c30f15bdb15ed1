"""GPU property tests at the full size of the 256^3 dam break (33.5M particles, 8^3 blocks, multigrid
PCG to 1e-3 per solve with FP64 iterative refinement to a 1e-5 relative FP64 residual), where the
oracle is too slow to step alongside.  They check what the size does not change:

  * the count field (FLIPMethod.cu:239-241, one per particle at each in-bounds corner) sums exactly to
    the number of in-bounds corners of the particles that entered the step;
  * the mass field (Rasterization.cu:121-190: trilinear weight + 1e-5 per corner) has the total and
    the first moments (sum of mass times cell centre, per axis) of a float64 scatter of the same
    particles on the CPU, to 1e-5 relative (dropping the 1e-5 term moves the total by 8e-5, a corner
    weight given to the wrong corner moves a moment);
  * every step meets the FP64 residual target, no particle leaves the closed box, the water falls
    (centre of mass lower after the steps);
  * two runs from the same particles give bit-identical particles and pressure.
"""
import hashlib

import numpy as np
import pytest

from vare_amd import _lib as L

pytestmark = pytest.mark.gpu

N, B, PPC = 256, 8, 8
DT = 1.0 / 240.0


def seed():
    pos = L.seed_box(N, N, N, 0.5 * N, 0.5 * N, float(N), PPC, 1)
    return pos, np.zeros_like(pos)


def handle(pos, vel):
    prm = L.default_params(ext_force=(0.0, -9.8, 0.0), wall_level=1e30, voxel_size=1.0 / N,
                           solver_mode=L.SOLVER_RELATIVE, cg_rel_tol=1e-3, refine_tol=1e-5, refine_max=16,
                           max_iterations=20000, preconditioner=L.PRECOND_MG)
    h = L.FlipHandle(L.grid_desc(N, N, N, block_size=B), prm, particle_capacity=len(pos),
                     block_capacity=(N // B) ** 3)
    h.set_particles(pos, vel)
    return h


def cpu_corner_sums(pos, n, chunk=1 << 22):
    """Per particle inside [0,n]^3: the base cell c = max((int)(p-0.5), 0) per axis and the corners c, c+1
    that are <= n-1.  Returns (number of in-bounds corners, sum of (weight + 1e-5) over them, and per
    axis the sum of (weight + 1e-5) * (corner + 0.5)), all in float64."""
    cnt = 0
    tot = 0.0
    mom = np.zeros(3)
    for a in range(0, len(pos), chunk):
        p = pos[a:a + chunk].astype(np.float64)
        inside = np.all((p >= 0.0) & (p <= n), axis=1)
        p = p[inside]
        c = np.maximum(np.floor(p - 0.5), 0.0)                     # (int) truncation, then max(., 0)
        w = []
        ok = []
        for d in (0, 1):
            w.append(np.maximum(0.0, 1.0 - np.abs(p - (c + d + 0.5))))
            ok.append(c + d <= n - 1)
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    m = ok[dx][:, 0] & ok[dy][:, 1] & ok[dz][:, 2]
                    wt = w[dx][:, 0] * w[dy][:, 1] * w[dz][:, 2] + 1e-5
                    wt = np.where(m, wt, 0.0)
                    cnt += int(m.sum())
                    tot += float(wt.sum())
                    for ax, dd in enumerate((dx, dy, dz)):
                        mom[ax] += float((wt * (c[:, ax] + dd + 0.5)).sum())
    return cnt, tot, mom


def field_sums(h):
    _, coords = h.block_table()
    cnt = h.cell_field(L.FIELD_COUNT).astype(np.float64).reshape(-1, B, B, B)        # [block][z][y][x]
    mass = h.cell_field(L.FIELD_MASS).astype(np.float64).reshape(-1, B, B, B)
    loc = np.arange(B, dtype=np.float64)
    mom = np.zeros(3)
    for ax, axis_shape in enumerate(((1, 1, 1, B), (1, 1, B, 1), (1, B, 1, 1))):
        centre = (coords[:, ax].astype(np.float64) * B).reshape(-1, 1, 1, 1) + loc.reshape(axis_shape) + 0.5
        mom[ax] = float((mass * centre).sum())
    return float(cnt.sum()), float(mass.sum()), mom


def test_dam_break_256_properties():
    pos, vel = seed()
    h = handle(pos, vel)
    n0 = len(pos)
    print("256^3 dam break: %d particles" % n0, flush=True)
    y0 = float(pos[:, 1].astype(np.float64).mean())
    for step in range(12):
        check_fields = step in (0, 10)
        if check_fields:
            before = h.get_particles()["pos"]
        h.step(DT)
        st = h.stats()
        print("step %d: %d PCG iterations, %d refinement rounds, FP64 rel. residual %.3e" %
              (step, st["cg_iterations"], st["refine_steps"], st["true_rel_residual"]), flush=True)
        assert st["converged"] and st["true_rel_residual"] <= 1e-5, st
        assert st["num_removed"] == 0 and st["num_particles"] == n0, st
        if check_fields:
            ccnt, ctot, cmom = cpu_corner_sums(before, N)
            gcnt, gtot, gmom = field_sums(h)
            print("  count %d vs %d, mass %.9e vs %.9e" % (gcnt, ccnt, gtot, ctot), flush=True)
            assert gcnt == ccnt
            assert abs(gtot - ctot) <= 1e-5 * ctot
            np.testing.assert_allclose(gmom, cmom, rtol=1e-5)
    y1 = float(h.get_particles()["pos"][:, 1].astype(np.float64).mean())
    print("centre of mass y: %.4f -> %.4f" % (y0, y1), flush=True)
    assert y1 < y0
    h.close()


def digest(h):
    p = h.get_particles()
    out = [hashlib.sha1(np.ascontiguousarray(p[k]).tobytes()).hexdigest() for k in ("pos", "vel", "adj")]
    out.append(hashlib.sha1(h.cell_field(L.FIELD_PRESSURE).tobytes()).hexdigest())
    return out


def test_dam_break_256_bit_reproducible():
    pos, vel = seed()
    runs = []
    for run in range(2):
        h = handle(pos, vel)
        its = []
        for _ in range(8):
            h.step(DT)
            its.append(h.stats()["cg_iterations"])
        runs.append((digest(h), its))
        print("run %d: PCG iterations %s" % (run, its), flush=True)
        h.close()
    assert runs[0] == runs[1]
