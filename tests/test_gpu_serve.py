"""Request server (include/peec_serve.h): each answer is the same when its
request runs alone and when it runs in a mix of overlapping requests under a
tight HBM budget (engines evicted and recreated), factors are reused for
repeated (mesh, frequency) points, and the answers agree with a standalone
engine and with the FP64 path."""
import numpy as np
import pytest

from peeeeeeeeeeeeeeeeeeec_amd import lib as plib
from peeeeeeeeeeeeeeeeeeec_amd import meshgen

pytestmark = pytest.mark.gpu


def _models():
    return [plib.Model(*meshgen.plate(8, 8)),
            plib.Model(*meshgen.box_by_size(0.6, 0.4, 0.3, 0.1)),
            plib.Model(*meshgen.plate(12, 6, 1.5, 0.75))]


def _requests(models):
    rng = np.random.default_rng(11)
    reqs = []
    for k in range(14):
        m = int(rng.integers(len(models)))
        nv = models[m].nv
        # repeated points (same mesh, frequency, mode) exercise the factor cache
        f = [1e6, 3e7, 1e8][k % 3] if k < 9 else float(10 ** rng.uniform(6, 8))
        nrhs = int(rng.integers(1, 5))
        ports = [tuple(int(x) for x in rng.choice(nv, 2, replace=False)) for _ in range(nrhs)]
        reqs.append((m, f, ports, plib.Server.MIXED if k % 4 else plib.Server.FP64))
    return reqs


def _run(models, reqs, slots, budget, together):
    sv = plib.Server(slots=slots, hbm_budget_bytes=budget, retarded=True)
    ids = [sv.add_mesh(m) for m in models]
    out = []
    if together:
        tickets = [sv.submit(ids[m], f, ports, mode) for m, f, ports, mode in reqs]
        out = [sv.wait(t) for t in tickets]
    else:
        for m, f, ports, mode in reqs:
            out.append(sv.wait(sv.submit(ids[m], f, ports, mode)))
    mem, n_eng = sv.memory()
    sv.close()
    return out, mem, n_eng


def test_server_results_alone_and_in_the_mix(require_gpu):
    models = _models()
    reqs = _requests(models)
    alone, _, _ = _run(models, reqs, slots=1, budget=0, together=False)
    # budget for about two engines at once: the mix evicts and recreates
    budget = 2 * max(plib.Server.engine_bytes(m.ne, m.nv, True) for m in models)
    mixed, mem, n_eng = _run(models, reqs, slots=4, budget=budget, together=True)
    assert mem <= budget
    for (st_a, za, ra), (st_b, zb, rb), (m, f, ports, mode) in zip(alone, mixed, reqs):
        assert st_a == 0 and st_b == 0, (m, f, mode)
        assert ra.report.backward_error <= 1e-10 and rb.report.backward_error <= 1e-10
        assert 0.0 <= ra.queue_s <= ra.latency_s
        np.testing.assert_allclose(zb, za, rtol=1e-10, atol=0)
    # the alone run reuses the factor of a repeated point on the same engine
    assert any(r.report.factor_cached for _, _, r in alone)
    # each answer against a standalone engine at that point (mixed or FP64)
    for i, (m, f, ports, mode) in enumerate(reqs[:5]):
        eng = plib.Engine(models[m], retarded=True)
        st, z, _, _, _ = eng.solve(f, ports, mode, want_solution=False)
        eng.close()
        assert st == 0
        np.testing.assert_allclose(alone[i][1], z, rtol=1e-10, atol=0)


def test_server_mixed_answers_against_fp64(require_gpu):
    models = _models()[:2]
    sv = plib.Server(slots=3, retarded=True)
    ids = [sv.add_mesh(m) for m in models]
    cases = [(0, 2e6, [(0, 80), (3, 40)]), (1, 5e7, [(0, models[1].nv - 1)]), (0, 9e7, [(10, 70)])]
    t_mix = [sv.submit(ids[m], f, p, plib.Server.MIXED) for m, f, p in cases]
    t_64 = [sv.submit(ids[m], f, p, plib.Server.FP64) for m, f, p in cases]
    for a, b in zip(t_mix, t_64):
        st1, z1, r1 = sv.wait(a)
        st2, z2, r2 = sv.wait(b)
        assert st1 == 0 and st2 == 0
        assert np.max(np.abs(z1 - z2) / np.abs(z2)) < 1e-9
    sv.close()


def test_server_rejects_bad_requests(require_gpu):
    model = _models()[0]
    sv = plib.Server(slots=1)
    mid = sv.add_mesh(model)
    with pytest.raises(RuntimeError):
        sv.submit(mid, 1e6, [(0, model.nv)])  # node out of range
    with pytest.raises(RuntimeError):
        sv.submit(mid + 1, 1e6, [(0, 1)])  # unknown mesh
    with pytest.raises(RuntimeError):
        sv.submit(mid, -1.0, [(0, 1)])
    sv.close()


def test_server_batches_queued_mixed_requests(require_gpu):
    """Continuous batching: mixed requests of one mesh that wait in the queue
    while the worker is busy run as one batched factor; each answer equals a
    standalone engine solving the same requests one by one in the same row
    order (set by an FP64 request first) and block width."""
    xyz, quads = meshgen.box(1.0, 0.6, 0.6, 10, 6, 6)
    nv = len(xyz)
    reqs = [(2e6, [(0, nv - 1)]), (7e6, [(1, nv - 2), (4, 50)]), (2e7, [(0, nv - 1)]),
            (6e7, [(7, nv - 9), (0, 30), (2, 60)]), (1.2e8, [(0, nv - 1)]), (1.9e8, [(20, 300)])]
    opts = dict(retarded=True, block_size=128)
    with plib.Model(xyz, quads) as m:
        sv = plib.Server(slots=1, max_batch=8, **opts)
        mid = sv.add_mesh(m)
        t0 = sv.submit(mid, 5e5, [(0, nv - 1)], plib.Server.FP64)
        tickets = [sv.submit(mid, f, p, plib.Server.MIXED) for f, p in reqs]
        st0, z0, _ = sv.wait(t0)
        out = [sv.wait(t) for t in tickets]
        sv.close()
        eng = plib.Engine(m, smooth_moments=0, **opts)
        s64, z64, _, _, _ = eng.solve(5e5, [(0, nv - 1)], plib.Engine.FP64, want_solution=False)
        single = [eng.solve(f, p, plib.Engine.MIXED, want_solution=False) for f, p in reqs]
        eng.close()
    assert st0 == 0 and s64 == 0
    np.testing.assert_allclose(z0, z64, rtol=1e-13, atol=0)
    sizes = [r.report.batch_size for _, _, r in out]
    print(f"batch sizes {sizes}", flush=True)
    assert max(sizes) >= 2
    for (st, z, r), (s1, z1, _, _, r1), (f, _) in zip(out, single, reqs):
        assert st == 0 and s1 == 0, f
        assert r.report.backward_error <= 1e-10
        d = np.max(np.abs(z - z1) / np.abs(z1))
        print(f"f={f:g} batch {r.report.batch_size} vs single {d:.2e}", flush=True)
        assert d <= 1e-13
