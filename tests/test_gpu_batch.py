"""Mixed-precision points of one mesh factored together in shared launches
(peec_engine_solve_batch): every batch member's impedances equal those of
the same point solved alone in the same row order and block width, and stay
within 1e-10 of the FP64 LU."""
from __future__ import annotations

import numpy as np
import pytest

from peeeeeeeeeeeeeeeeeeec_amd import lib as plib
from peeeeeeeeeeeeeeeeeeec_amd import meshgen

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("block", [128, 64])
def test_batch_equals_single_points(require_gpu, block):
    """Box N = 1538, retarded, five points: the first alone (it sets the row
    order), four in one batched factor; then a second batch of three points
    with different port sets. Each member against the single-point engine
    (same options) and against FP64."""
    xyz, quads = meshgen.box(1.0, 0.6, 0.6, 10, 6, 6)
    nv = len(xyz)
    port = (0, nv - 1)
    freqs = [1e6, 3e6, 1e7, 3e7, 1e8]
    more = [2e6, 5e7, 1.5e8]
    more_ports = [[port], [(1, nv - 2), (5, 40)], [(3, nv - 7), (0, 9), (12, 100)]]
    with plib.Model(xyz, quads) as m:
        print(f"[batch] N = {m.ne + m.nv}, block {block}", flush=True)
        ref = plib.Engine(m, retarded=True, block_size=block)
        single = [ref.solve(f, [port], plib.Engine.MIXED, want_solution=False) for f in freqs]
        single_more = [ref.solve(f, p, plib.Engine.MIXED, want_solution=False) for f, p in zip(more, more_ports)]
        fp64 = [ref.solve(f, [port], plib.Engine.FP64, want_solution=False)[1] for f in freqs]
        fp64_more = [ref.solve(f, p, plib.Engine.FP64, want_solution=False)[1] for f, p in zip(more, more_ports)]
        ref.close()
        eng = plib.Engine(m, retarded=True, block_size=block)
        code, st, zs, reps = eng.solve_batch(freqs, [port])
        code2, st2, zs2, reps2 = eng.solve_batch(more, more_ports)
        eng.close()
    assert code == 0 and code2 == 0 and not st.any() and not st2.any()
    assert reps[0].batch_size == 1 and all(r.batch_size == 4 for r in reps[1:])
    assert all(r.batch_size == 3 for r in reps2)
    for f, z, r, (s1, z1, _, _, r1), z64 in zip(freqs + more, zs + zs2, reps + reps2, single + single_more,
                                                 fp64 + fp64_more):
        d = np.max(np.abs(z - z1) / np.abs(z1))
        e64 = np.max(np.abs(z - z64) / np.abs(z64))
        print(f"f={f:g} batch {r.batch_size} reuse {r.pivot_reuse} steps {r.refinement_steps}/{r1.refinement_steps} "
              f"eta {r.backward_error:.2e} vs single {d:.2e} vs fp64 {e64:.2e} factor {r.ms_factor:.2f} ms",
              flush=True)
        assert s1 == 0
        assert r.backward_error <= 1e-10
        assert r.refinement_steps == r1.refinement_steps
        assert d <= 1e-14
        assert e64 <= 1e-10


def test_batch_shape_autotuner(require_gpu):
    """With block_size = 0 the batched factors are autotuned apart from the
    single ones: successive batches try block widths 128 / 64 with update
    tiles of 64 / 32 rows, then keep the fastest per member; every answer
    stays within 1e-10 of FP64."""
    xyz, quads = meshgen.box(1.0, 0.6, 0.6, 10, 6, 6)
    port = (0, len(xyz) - 1)
    batches = [[2e6 * (1.3 ** (3 * b + i)) for i in range(3)] for b in range(5)]
    with plib.Model(xyz, quads) as m:
        eng = plib.Engine(m, retarded=True)
        st0, _, _, _, r0 = eng.solve(1e6, [port], plib.Engine.MIXED, want_solution=False)
        assert st0 == 0 and r0.pivot_reuse == 0
        shapes, answers = [], []
        for fs in batches:
            code, st, zs, reps = eng.solve_batch(fs, [port])
            assert code == 0 and not st.any()
            assert all(r.batch_size == 3 for r in reps)
            shapes.append({(r.block_size, r.update_tile_rows) for r in reps})
            answers += [(f, z[0], r) for f, z, r in zip(fs, zs, reps)]
        for f, z, r in answers:
            _, z64, _, _, _ = eng.solve(f, [port], plib.Engine.FP64, want_solution=False)
            assert abs(z - z64[0]) <= 1e-10 * abs(z64[0]) and r.backward_error <= 1e-10
        eng.close()
    print(f"batch shapes {shapes}", flush=True)
    assert all(len(s) == 1 for s in shapes)
    tried = [s.pop() for s in shapes]
    assert tried[:4] == [(128, 64), (128, 32), (64, 64), (64, 32)]
    assert tried[4] in tried[:4]
