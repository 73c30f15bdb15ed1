"""Mixed precision against the FP64 oracle: the pool + sphere scene (the 512^3 benchmark scene) at
64^3 with FP16 particle velocities, FP32 multigrid PCG and FP64 iterative refinement to 1e-8.

Errors reported (printed) and bounded:
  * pressure and projected grid velocity: relative L2 against the FP64 oracle;
  * particle velocities after the step: relative L2 (dominated by the FP16 storage, 2^-11);
  * post-projection divergence ratio D = |div_c v_post| / |div_c v_pre| with the reference's
    central-difference divergence (the right-hand side of FluidProjection::buildLinearSystem):
    |D_mixed - D_fp64| <= 1e-5.
"""
import numpy as np
import pytest

from oracle import oracle as O
import vare_amd._lib as L

from test_gpu_parity import DT, rel

pytestmark = pytest.mark.gpu


def _div_norm(o, comp):
    """|b(comp)| with the oracle's divergence rule on its current classification."""
    o.set_field(O.FIELD_COMP, comp)
    o.stage(2)
    return float(np.linalg.norm(o.field(O.FIELD_B)))


@pytest.mark.parametrize("apply32", [False, True])
def test_pool_sphere_fp16_refined_vs_fp64_oracle(apply32):
    """apply32: the refined FP64 pressure is rounded to FP32 before the air extrapolation and the
    gradient (VARE_FLAG_FP32_APPLY); otherwise both run in FP64 on the refined pressure."""
    n, B = 64, 8
    pos = L.seed_pool_sphere(n, n, n, 0.3 * n, (0.5 * n, 0.7 * n, 0.5 * n), 0.2 * n, 8, 1)
    kw = dict(ext_force=(0.0, -9.8, 0.0), wall_level=1e30, voxel_size=1.0 / n, solver_mode=L.SOLVER_RELATIVE,
              preconditioner=L.PRECOND_MG, cg_rel_tol=1e-3, refine_tol=1e-8, refine_max=16, max_iterations=4000)
    g = L.grid_desc(n, n, n, block_size=B)
    # give the scene motion first (sphere falling), then restart both sides from that state
    h0 = L.FlipHandle(g, L.default_params(velocity_fp16=1, **kw))
    h0.set_particles(pos, np.zeros_like(pos))
    for _ in range(6):
        h0.step(DT)
    st = h0.get_particles()
    p0, v0 = st["pos"], st["vel"]                    # v0 holds FP16 values exactly
    print("case pool+sphere 64^3 np=%d" % p0.shape[0], flush=True)

    flags = L.FLAG_REPORT_DIVERGENCE | (L.FLAG_FP32_APPLY if apply32 else 0)
    h = L.FlipHandle(g, L.default_params(velocity_fp16=1, flags=flags, **kw))
    h.set_particles(p0, v0)
    o64 = O.Oracle((n, n, n), B=B, prm=O.params(ext_force=(0.0, -9.8, 0.0), wall_level=1e30, voxel_size=1.0 / n,
                                                solver_mode=1, cg_rel_tol=1e-12, max_iterations=20000,
                                                preconditioner=1), fp64=True)
    o64.set_particles(p0, v0)
    for s in range(3):
        h.stage(s, DT)
        o64.stage(s, DT)
    div_pre = float(np.linalg.norm(o64.field(O.FIELD_B)))
    h.stage(3, DT)
    o64.stage(3, DT)
    stats = h.stats()
    assert stats["true_rel_residual"] <= 1e-8 * 1.0001 and stats["refine_steps"] >= 1
    e_p = rel(h.cell_field(L.FIELD_PRESSURE), o64.field(O.FIELD_PRES))
    comp_gpu = h.cell_field(L.FIELD_COMP).astype(np.float64)
    comp_o = o64.field(O.FIELD_COMP)
    e_v = rel(comp_gpu, comp_o)
    # divergence of both projected fields under the same rule and classification
    d_o = _div_norm(o64, comp_o) / div_pre
    d_g = _div_norm(o64, comp_gpu) / div_pre
    o64.set_field(O.FIELD_COMP, comp_o)
    h.stage(4, DT)
    o64.stage(4, DT)
    e_pv = rel(h.get_particles()["vel"], o64.get_particles()["vel"])
    e_pp = rel(h.get_particles()["pos"], o64.get_particles()["pos"])
    print("apply %s: errors vs FP64 oracle: pressure %.2e  grid velocity %.2e  particle velocity %.2e  "
          "position %.2e  D_mixed %.6e (on the GPU %.6e)  D_fp64 %.6e  |dD| %.2e  refine %d  FP64 residual %.2e"
          % ("FP32" if apply32 else "FP64", e_p, e_v, e_pv, e_pp, d_g, stats["div_ratio"], d_o, abs(d_g - d_o),
             stats["refine_steps"], stats["true_rel_residual"]), flush=True)
    # the step's own report of D (FP32 divergence, FP64 sums) is the oracle's D of the same field
    assert abs(stats["div_ratio"] - d_g) <= 1e-6 * d_g
    assert e_p < 1e-5
    assert e_v < 1e-5
    assert abs(d_g - d_o) <= 1e-5
    assert e_pv < 2e-3          # FP16 storage of the particle velocity (unit roundoff 2^-11)
    assert e_pp < 1e-6
