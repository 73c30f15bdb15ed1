"""GPU tests of the continuous-batching server (vare_server_*): admission against the HBM budget and
the slot count, FIFO order, batched rounds (every resident tank stepped with shared launches), and the
same result for a tank solved alone and in a mix."""
import numpy as np
import pytest

from vare_amd import _lib as L

pytestmark = pytest.mark.gpu

DT = 1.0 / 240.0


def tank(n, frac_x, frac_y, ppc, seed):
    pos = L.seed_box(n, n, n, frac_x * n, frac_y * n, float(n), ppc, seed)
    return pos, np.zeros_like(pos)


def params(n):
    return L.default_params(ext_force=(0.0, -9.8, 0.0), wall_level=1e30, voxel_size=1.0 / n, solver_mode=1,
                            cg_rel_tol=1e-5, preconditioner=L.PRECOND_MG)


def rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(np.asarray(b, np.float64)), 1e-30)


def test_alone_vs_mix_same_result():
    n0 = 48
    pos0, vel0 = tank(n0, 0.4, 0.6, 6, 1)
    g0 = L.grid_desc(n0, n0, n0, block_size=8)
    alone = L.Server(8 << 30, 1)
    a = alone.submit(g0, params(n0), pos0, vel0, steps=6, dt=DT)
    alone.run()
    pa, va, sa = alone.result(a)
    assert sa["status"] == 0 and sa["steps_done"] == 6

    mix = L.Server(8 << 30, 4)
    ids = []
    for k, (n, fx, fy, ppc) in enumerate([(32, 0.5, 0.3, 4), (64, 0.3, 0.5, 5), (40, 0.6, 0.6, 8)]):
        p, v = tank(n, fx, fy, ppc, 10 + k)
        ids.append(mix.submit(L.grid_desc(n, n, n, block_size=8), params(n), p, v, steps=6, dt=DT,
                              arrival_s=0.0))
    b = mix.submit(g0, params(n0), pos0, vel0, steps=6, dt=DT, arrival_s=0.01)
    mix.run()
    pb, vb, sb = mix.result(b)
    st = mix.stats()
    print("mix stats", st, flush=True)
    assert st["max_resident"] >= 2                       # work overlapped on the device
    assert st["max_batch"] >= 2 and st["rounds"] >= 6    # stepped together in shared launches
    assert pb.shape == pa.shape
    # same tank alone and in the mix: every kernel is deterministic and a batched launch gives each
    # tank exactly its own launch's result, so the two are bit-identical
    assert np.array_equal(pb, pa) and np.array_equal(vb, va)


def test_batched_and_stream_serving_agree():
    # the same four requests served with batched rounds and with one stream per tank: identical results
    reqs = []
    for k, (n, fx, fy, ppc) in enumerate([(32, 0.5, 0.3, 4), (48, 0.3, 0.5, 5), (40, 0.6, 0.6, 8), (56, 0.4, 0.4, 4)]):
        p, v = tank(n, fx, fy, ppc, 40 + k)
        reqs.append((L.grid_desc(n, n, n, block_size=8), params(n), p, v, 3 + k, 0.002 * k))
    out = {}
    for batching in (True, False):
        srv = L.Server(8 << 30, 4, batching=batching)
        ids = [srv.submit(g, prm, p, v, steps=st, dt=DT, arrival_s=a) for g, prm, p, v, st, a in reqs]
        srv.run()
        out[batching] = [srv.result(i) for i in ids]
        st = srv.stats()
        print("batching=%s stats %s" % (batching, st), flush=True)
        assert (st["rounds"] > 0) == batching
    for (p1, v1, s1), (p0, v0, s0) in zip(out[True], out[False]):
        assert s1["status"] == 0 and s0["status"] == 0 and s1["steps_done"] == s0["steps_done"]
        assert np.array_equal(p1, p0) and np.array_equal(v1, v0)


def test_admission_budget_and_fifo():
    # budget for two 48^3 tanks at a time: never more than two resident, admitted in arrival order
    n = 48
    g = L.grid_desc(n, n, n, block_size=8)
    pos, vel = tank(n, 0.4, 0.4, 4, 3)
    per = L.Server.request_bytes(g, params(n), pos.shape[0])
    srv = L.Server(2 * per + 1024, 8)
    ids = [srv.submit(g, params(n), pos, vel, steps=3, dt=DT, arrival_s=0.002 * k) for k in range(5)]
    srv.run()
    st = srv.stats()
    assert st["max_resident"] <= 2 and st["requests"] == 5 and st["peak_bytes"] <= 2 * per + 1024
    admits = [srv.result(i)[2]["admit_s"] for i in ids]
    assert admits == sorted(admits)
    assert st["step_p99_ms"] >= st["step_p50_ms"] > 0


def test_request_over_budget_rejected():
    n = 64
    g = L.grid_desc(n, n, n, block_size=8)
    pos, vel = tank(n, 0.5, 0.5, 8, 4)
    srv = L.Server(1 << 20, 2)
    with pytest.raises(L.VareError) as e:
        srv.submit(g, params(n), pos, vel, steps=1, dt=DT)
    assert e.value.code == L.VARE_EBUDGET


def test_pooled_handle_reuse_same_result():
    # one slot: the second tank (same grid) is served by the first tank's pooled device state and must
    # give exactly what a fresh handle gives (every kernel of the step is deterministic)
    n = 40
    g = L.grid_desc(n, n, n, block_size=8)
    pa, va = tank(n, 0.6, 0.5, 6, 21)
    pb, vb = tank(n, 0.3, 0.6, 5, 22)
    fresh = L.Server(8 << 30, 1)
    b0 = fresh.submit(g, params(n), pb, vb, steps=4, dt=DT)
    fresh.run()
    p0, v0, s0 = fresh.result(b0)
    srv = L.Server(8 << 30, 1)
    srv.submit(g, params(n), pa, va, steps=4, dt=DT, arrival_s=0.0)
    b1 = srv.submit(g, params(n), pb, vb, steps=4, dt=DT, arrival_s=0.001)
    srv.run()
    st = srv.stats()
    assert st["handles_created"] == 1 and st["handles_reused"] == 1, st
    p1, v1, s1 = srv.result(b1)
    assert s0["status"] == 0 and s1["status"] == 0
    assert np.array_equal(p1, p0) and np.array_equal(v1, v0)


def test_pool_eviction_under_budget():
    # one slot; tanks of grids A, B, A in turn.  With room for only one device state the pooled A is
    # destroyed to admit B and B to admit the second A (three states created, none reused, the
    # reserved bytes never above the budget); with room for both, the second A reuses the first's.
    na, nb = 40, 48
    ga, gb = L.grid_desc(na, na, na, block_size=8), L.grid_desc(nb, nb, nb, block_size=8)
    pa, va = tank(na, 0.5, 0.5, 4, 31)
    pb, vb = tank(nb, 0.5, 0.5, 4, 32)
    bya = L.Server.request_bytes(ga, params(na), pa.shape[0])
    byb = L.Server.request_bytes(gb, params(nb), pb.shape[0])
    for budget, created, reused in ((max(bya, byb) + 1024, 3, 0), (bya + byb + 1024, 2, 1)):
        srv = L.Server(budget, 1)
        ids = [srv.submit(ga, params(na), pa, va, steps=2, dt=DT, arrival_s=0.0),
               srv.submit(gb, params(nb), pb, vb, steps=2, dt=DT, arrival_s=0.001),
               srv.submit(ga, params(na), pa, va, steps=2, dt=DT, arrival_s=0.002)]
        srv.run()
        st = srv.stats()
        assert (st["handles_created"], st["handles_reused"]) == (created, reused), st
        assert st["requests"] == 3 and st["peak_bytes"] <= budget
        r0, r2 = srv.result(ids[0]), srv.result(ids[2])
        assert r0[2]["status"] == 0 and r2[2]["status"] == 0
        assert np.array_equal(r0[0], r2[0]) and np.array_equal(r0[1], r2[1])   # same tank, same result


def test_served_tank_matches_fp64_oracle():
    # one tank through the server (MG-PCG to 1e-5 relative) against the FP64 oracle's step on the same
    # input; the bar is the one of the stage-wise parity test: within max(2 x the FP32 oracle's own
    # error, 1e-5) of the FP64 oracle, positions within 1e-6
    from oracle import oracle as O
    n = 48
    pos = O.seed_box(n, n, n, 0.5 * n, 0.6 * n, float(n), 6, 51)
    rng = np.random.default_rng(51)
    vel = (rng.standard_normal(pos.shape) * 10).astype(np.float32)
    g = L.grid_desc(n, n, n, block_size=8)
    srv = L.Server(8 << 30, 2)
    rid = srv.submit(g, params(n), pos, vel, steps=1, dt=DT)
    srv.run()
    ps, vs, st = srv.result(rid)
    assert st["status"] == 0 and st["steps_done"] == 1
    kw = dict(ext_force=(0.0, -9.8, 0.0), wall_level=1e30, voxel_size=1.0 / n, solver_mode=1, max_iterations=20000,
              preconditioner=1)
    o64 = O.Oracle((n, n, n), B=8, prm=O.params(cg_rel_tol=1e-12, **kw), fp64=True)
    o32 = O.Oracle((n, n, n), B=8, prm=O.params(cg_rel_tol=1e-5, **kw), fp64=False)
    for o in (o64, o32):
        o.set_particles(pos, vel)
        o.step(DT)
    r64, r32 = o64.get_particles(), o32.get_particles()
    assert ps.shape[0] == r64["pos"].shape[0]
    ev, ev32 = rel(vs, r64["vel"]), rel(r32["vel"], r64["vel"])
    print("served tank vs FP64 oracle: velocity %.3g (FP32 oracle %.3g), position %.3g" % (
        ev, ev32, rel(ps, r64["pos"])), flush=True)
    assert rel(ps, r64["pos"]) < 1e-6
    assert ev < max(2 * ev32, 1e-5)


def test_poisson_stream_256_properties():
    # the served-batch configuration at its full request count (one step each): 256 tanks with edges
    # log-uniform in 32..128 (multiples of 8), 4..8 particles per cell, water fractions U(0.2, 0.7),
    # Poisson arrivals, a 200 GB budget and 16 slots.  Every request succeeds, admission follows
    # arrival order, the reserved bytes stay within the budget, and sampled tanks give exactly what a
    # handle of their own gives.
    rng = np.random.default_rng(7)
    reqs = []
    for k in range(256):
        edge = int(8 * round(np.exp(rng.uniform(np.log(32), np.log(128))) / 8))
        ppc = int(rng.integers(4, 9))
        fx, fy = rng.uniform(0.2, 0.7, 2)
        pos = L.seed_box(edge, edge, edge, fx * edge, fy * edge, float(edge), ppc, 500 + k)
        reqs.append((edge, pos))
    arrivals = np.cumsum(rng.exponential(0.002, 256))
    arrivals -= arrivals[0]
    budget = 200 << 30
    srv = L.Server(budget, 16)
    ids = [srv.submit(L.grid_desc(e, e, e, block_size=8), params(e), p, np.zeros_like(p), steps=1, dt=DT,
                      arrival_s=float(a)) for (e, p), a in zip(reqs, arrivals)]
    srv.run()
    st = srv.stats()
    print("stream of 256: %s" % st, flush=True)
    res = [srv.result(i) for i in ids]
    assert all(r[2]["status"] == 0 and r[2]["steps_done"] == 1 for r in res)
    assert st["requests"] == 256 and st["peak_bytes"] <= budget and st["max_resident"] <= 16
    assert st["max_batch"] >= 2                          # arrivals overlapped: tanks stepped together
    admits = [r[2]["admit_s"] for r in res]                # ids are in arrival order
    assert admits == sorted(admits)
    assert all(r[2]["admit_s"] >= r[2]["arrival_s"] for r in res)
    for k in (0, 37, 101, 180, 255):
        e, p = reqs[k]
        h = L.FlipHandle(L.grid_desc(e, e, e, block_size=8), params(e))
        h.set_particles(p, np.zeros_like(p))
        h.step(DT)
        d = h.get_particles()
        assert np.array_equal(res[k][0], d["pos"]) and np.array_equal(res[k][1], d["vel"]), k
        h.close()
