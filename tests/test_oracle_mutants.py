"""Mutation check of the oracle's pins (CPU only).

Each mutant is a copy of oracle/flip_oracle.cpp with ONE plausible restatement mistake planted in the
boundary rules of FluidProjection::buildLinearSystem (base/source/FluidProjection.cu:206-244) or
FluidProjection::solve (:297-378), in the frame post-process (FLIPMethod.cu:300-405) or in the solver's
dot product (BLAS.cu:98-104): a flipped
sign, a swapped neighbour index or component, a dropped extrapolation or normalisation, a wrong
divisor or exponent.  The copy is compiled into a temporary directory and the pins
(tests/test_oracle_boundary.py, tests/test_oracle_post.py) are run against it (VARE_ORACLE_LIB);
every mutant must make its pin file fail, and the unmodified source compiled the same way must pass.  The mistakes live only in the
temporary copies: the oracle itself is never changed.
"""
import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "flip_oracle.cpp")
PINS = os.path.join(ROOT, "tests", "test_oracle_boundary.py")

# (name, exact text in flip_oracle.cpp, replacement); each text occurs once
MUTANTS = [
    # FluidProjection.cu:220-235 (b)
    ("b: air ghost sign", "t_i1 == kAir ? R(2) * v.x - i0_j_k", "t_i1 == kAir ? R(2) * v.x + i0_j_k"),
    ("b: air ghost index", "t_j0 == kAir ? R(2) * v.y - i_j1_k", "t_j0 == kAir ? R(2) * v.y - i_j0_k"),
    ("b: low wall sign", "t_i0 == kWall ? Abs(v.x)", "t_i0 == kWall ? -Abs(v.x)"),
    ("b: high wall sign", "t_k1 == kWall ? -Abs(v.z)", "t_k1 == kWall ? Abs(v.z)"),
    ("b: domain-edge extrapolation dropped", "i == 0 ? R(2) * v.x - Vf(i + 1, j, k).x", "i == 0 ? v.x"),
    ("b: divergence sign", "const R d = -((v_i1 - v_i0)", "const R d = ((v_i1 - v_i0)"),
    ("b: face type at the domain edge",
     "((float)j + 0.5f <= wallLv + 1.f) ? (uint64_t)kWall : (uint64_t)kAir",
     "((float)j + 0.5f <= wallLv + 1.f) ? (uint64_t)kAir : (uint64_t)kWall"),
    # FluidProjection.cu:237-244 (diagonal)
    ("A: air does not raise the diagonal", "if (tn[q] == kAir) p += R(1);", "if (tn[q] == kAir) p += R(0);"),
    # FluidProjection.cu:314-335 (air-cell mean)
    ("p: air mean divisor", "pres[n] = press / (R)cnt;", "pres[n] = press / (R)(cnt + 1);"),
    ("p: air mean neighbour index", "{ press += P(i, j + 1, k); cnt++; }", "{ press += P(i + 1, j, k); cnt++; }"),
    # FluidProjection.cu:360-375 (gradient)
    ("grad: solid mirror sign", "t_i0 == kSolid ? R(2) * p - i1_j_k", "t_i0 == kSolid ? R(2) * p + i1_j_k"),
    ("grad: solid mirror index", "t_j1 == kSolid ? R(2) * p - i_j0_k", "t_j1 == kSolid ? R(2) * p - i_j1_k"),
    ("grad: air ghost sign", "t_i1 == kAir ? -i1_j_k", "t_i1 == kAir ? i1_j_k"),
    ("grad: domain-edge extrapolation dropped", "i == 0 ? R(2) * p - P(i + 1, j, k)", "i == 0 ? p"),
    ("grad: subtract sign", "comp[n].y -= (p_j1 - p_j0) * R(0.5);", "comp[n].y += (p_j1 - p_j0) * R(0.5);"),
    ("grad: face type at the domain edge",
     "((float)j + 0.5f <= wallLv + 1.f) ? (uint64_t)kSolid : (uint64_t)kAir",
     "((float)j + 0.5f <= wallLv + 1.f) ? (uint64_t)kAir : (uint64_t)kSolid"),
]

# FLIPMethod::postProcess (FLIPMethod.cu:300-405), pinned by tests/test_oracle_post.py
POST_PINS = os.path.join(ROOT, "tests", "test_oracle_post.py")
POST_MUTANTS = [
    ("splat: droplet added", "surf[n] = std::sqrt(ex * ex + ey * ey + ez * ez) - droplet;",
     "surf[n] = std::sqrt(ex * ex + ey * ey + ez * ez) + droplet;"),
    ("splat: kernel power", "const R w = Max(R(0), t * t * t);", "const R w = Max(R(0), t * t);"),
    ("redistance: upwind sides swapped", "P2(Max(dx0, R(0))) + P2(Min(dx1, R(0)))", "P2(Min(dx0, R(0))) + P2(Max(dx1, R(0)))"),
    ("redistance: update sign", "tmp[n] = v - dt * Max(sign, R(0)) * (gp - R(1))", "tmp[n] = v + dt * Max(sign, R(0)) * (gp - R(1))"),
    ("redistance: tank boundary below", "v_j0 = j == 0 ? ll - R(1) : F(i, j - 1, k);", "v_j0 = j == 0 ? ll + R(1) : F(i, j - 1, k);"),
    ("curl: y component reversed", "(z1.x - z0.x) * R(0.5) - (x1.z - x0.z) * R(0.5)", "(x1.z - x0.z) * R(0.5) - (z1.x - z0.x) * R(0.5)"),
    ("normal gradient not normalised", "out[n] = normalized(V3<R>{(i1 - i0) * R(0.5), (j1 - j0) * R(0.5), (k1 - k0) * R(0.5)});",
     "out[n] = V3<R>{(i1 - i0) * R(0.5), (j1 - j0) * R(0.5), (k1 - k0) * R(0.5)};"),
    ("divergence: wrong component", "const R j1 = j == EJ ? s.y * R(2) - F(i, j - 1, k).y : F(i, j + 1, k).y;",
     "const R j1 = j == EJ ? s.y * R(2) - F(i, j - 1, k).y : F(i, j + 1, k).x;"),
    ("vorticity: decay factor", "vort[n] -= vort[n] * dis;", "vort[n] *= dis;"),
    ("vorticity: threshold sign", "v = Max(R(0), v - R(0.2f));", "v = Max(R(0), v + R(0.2f));"),
]

# BLAS::dot (BLAS.cu:98-104), pinned by tests/test_oracle_solver.py
DOT_PINS = os.path.join(ROOT, "tests", "test_oracle_solver.py")
DOT_MUTANTS = [
    ("dot: a times a", "for (size_t n = 0; n < a.size(); ++n) s += (double)a[n] * (double)c[n];",
     "for (size_t n = 0; n < a.size(); ++n) s += (double)a[n] * (double)a[n];"),
    ("dot: last term dropped", "for (size_t n = 0; n < a.size(); ++n) s += (double)a[n] * (double)c[n];",
     "for (size_t n = 0; n + 1 < a.size(); ++n) s += (double)a[n] * (double)c[n];"),
    ("dot: accumulated in the solver's precision", "        double s = 0.0;\n#pragma omp parallel for reduction(+ : s)",
     "        R s = 0;\n#pragma omp parallel for reduction(+ : s)"),
]

FLAGS = ["-O1", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math"]


def _build_and_run(src_text, workdir, tag, pins=PINS):
    cpp = os.path.join(workdir, tag + ".cpp")
    so = os.path.join(workdir, tag + ".so")
    with open(cpp, "w") as f:
        f.write(src_text)
    b = subprocess.run([os.environ.get("CXX", "g++")] + FLAGS + ["-o", so, cpp], capture_output=True, text=True)
    if b.returncode != 0:
        return "build failed: " + b.stderr[-400:]
    env = dict(os.environ, VARE_ORACLE_LIB=so)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", pins], cwd=ROOT,
                       env=env, capture_output=True, text=True, timeout=300)
    return r.returncode


@pytest.mark.parametrize("mutants,pins", [(MUTANTS, PINS), (POST_MUTANTS, POST_PINS), (DOT_MUTANTS, DOT_PINS)],
                         ids=["boundary", "post", "dot"])
def test_every_planted_mistake_fails_a_pin(mutants, pins):
    src = open(SRC).read()
    for name, old, new in mutants:
        assert src.count(old) == 1, "mutant %r: text must occur exactly once" % name
    with tempfile.TemporaryDirectory() as d:
        jobs = [("control", src)] + [("m%02d" % q, src.replace(old, new)) for q, (_, old, new) in enumerate(mutants)]
        with ThreadPoolExecutor(max_workers=4) as ex:
            res = dict(zip([t for t, _ in jobs], ex.map(lambda j: _build_and_run(j[1], d, j[0], pins), jobs)))
    assert res["control"] == 0, "unmodified oracle must pass the pins: %r" % (res["control"],)
    survivors = [mutants[q][0] for q in range(len(mutants)) if res["m%02d" % q] != 1]
    assert not survivors, "mutants not caught (pytest status != 1): %s" % [
        (n, res["m%02d" % q]) for q, (n, _, _) in enumerate(mutants) if n in survivors]
