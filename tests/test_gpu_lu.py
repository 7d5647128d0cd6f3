"""GPU LU, drop-in harmonic path and the frequency-point engine against the
reference's known answers, golden vectors and the oracle."""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np
import pytest

from peeeeeeeeeeeeeeeeeeec_amd import ctypes_api as api
from peeeeeeeeeeeeeeeeeeec_amd import lib as plib
from peeeeeeeeeeeeeeeeeeec_amd import meshgen
from tests import checker as ck

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_mfma_layout_probe(require_gpu):
    """The accumulator layouts the GEMM kernels assume (lu.hip Mfma::row)."""
    out = np.zeros(1024)
    assert plib.load().peec_selftest_mfma_layout(out.ctypes.data_as(C.POINTER(C.c_double))) == 0
    A = np.array([[i + 16 * k + 1 for k in range(4)] for i in range(16)], float)
    D = np.zeros((16, 16))
    D[:, :4] = A
    D2 = np.tile(A[:, :1], (1, 16))
    rows = {"f64": lambda l, v: 4 * v + l // 16, "f32": lambda l, v: 4 * (l // 16) + v}
    for name, off in (("f64", 0), ("f32", 512)):
        d = out[off:off + 256].reshape(64, 4)
        d2 = out[off + 256:off + 512].reshape(64, 4)
        for lane in range(64):
            for v in range(4):
                i = rows[name](lane, v)
                assert d[lane, v] == D[i, lane % 16] and d2[lane, v] == D2[i, lane % 16], (name, lane, v)


def test_complex_lu_known_answer(require_gpu):
    """test_core.c:41-66 (needs a row swap), error < 1e-11."""
    A = np.array([[0, 2 + 1j], [1 - 1j, 3 + 0.5j]])
    x = np.array([1 + 2j, -0.5 + 0.75j])
    st, X, info = plib.complex_lu_factor_solve(A, A @ x)
    assert st == 0 and np.abs(X[0] - x).max() < 1e-11
    assert info["pivots"][0] == 1 and info["swap_count"] == 1


def test_complex_lu_singular(require_gpu):
    """test_core.c:328-337."""
    A = np.array([[1 + 1j, 2 + 2j], [2 + 2j, 4 + 4j]])
    st, _, _ = plib.complex_lu_factor_solve(A, np.ones(2))
    assert st == -2


def test_real_lu_known_answer(require_gpu):
    """test_core.c:17-39 and the determinant (lu.c:732-750)."""
    A = np.array([[0, 2, 1], [1, -2, -3], [2, 3, 1]], float)
    x = np.array([1, -2, 0.5])
    st, X, det = plib.real_lu_factor_solve(A, A @ x)
    assert st == 0 and np.abs(X[0] - x).max() < 1e-11
    assert abs(det - np.linalg.det(A)) < 1e-12 * abs(np.linalg.det(A))


@pytest.mark.parametrize("n", [7, 130, 515, 1100])
def test_complex_lu_random_against_oracle(require_gpu, n):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A *= rng.uniform(1e-3, 1e3, size=(n, 1))
    B = rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))
    st, X, info = plib.complex_lu_factor_solve(A, B)
    assert st == 0
    st_o, fac = ck.oracle_complex_lu(A)
    assert st_o == 0
    # equilibration is element-wise: the scales agree to the ulp of hypot
    np.testing.assert_allclose(info["row_scale"], fac[2], rtol=5e-16, atol=0)
    np.testing.assert_allclose(info["column_scale"], fac[3], rtol=1e-15, atol=0)
    for k in range(3):
        xo = ck.oracle_complex_solve(fac, B[k])
        assert np.abs(X[k] - xo).max() <= 1e-10 * np.abs(xo).max()
        eta = ck.oracle_backward_error(A, X[k], B[k], n // 2)[2]
        assert eta < 1e-13


@pytest.mark.parametrize("workgroups", [1, 3])
def test_complex_lu_solve_chain_with_few_workgroups(require_gpu, workgroups, monkeypatch):
    """The triangular-solve chain with fewer workgroups than 64-row blocks
    (each workgroup takes blocks w, w + G, ...), as on systems too large for
    one workgroup per block."""
    monkeypatch.setenv("PEEC_CHAIN_WORKGROUPS", str(workgroups))
    n = 515
    rng = np.random.default_rng(77)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    B = rng.standard_normal((10, n)) + 1j * rng.standard_normal((10, n))
    st, X, _ = plib.complex_lu_factor_solve(A, B)
    assert st == 0
    st_o, fac = ck.oracle_complex_lu(A)
    for k in range(10):
        xo = ck.oracle_complex_solve(fac, B[k])
        assert np.abs(X[k] - xo).max() <= 1e-10 * np.abs(xo).max()


@pytest.mark.parametrize("n", [70, 515])
def test_complex_lu_many_rhs_blocked_solve(require_gpu, n):
    """40 right-hand sides take the blocked solve (diagonal-block inverses +
    MFMA updates on the transposed right-hand sides) instead of the chain."""
    rng = np.random.default_rng(n + 1)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    B = rng.standard_normal((40, n)) + 1j * rng.standard_normal((40, n))
    st, X, _ = plib.complex_lu_factor_solve(A, B)
    assert st == 0
    st_o, fac = ck.oracle_complex_lu(A)
    for k in (0, 17, 39):
        xo = ck.oracle_complex_solve(fac, B[k])
        assert np.abs(X[k] - xo).max() <= 1e-10 * np.abs(xo).max()


def test_real_lu_many_rhs_blocked_solve(require_gpu):
    rng = np.random.default_rng(9)
    n = 300
    A = rng.standard_normal((n, n))
    B = rng.standard_normal((20, n))
    st, X, _ = plib.real_lu_factor_solve(A, B)
    assert st == 0
    ref = np.linalg.solve(A, B.T).T
    assert np.abs(X - ref).max() <= 1e-10 * np.abs(ref).max()


def test_engine_many_ports_mixed_vs_fp64(require_gpu):
    """A mixed point with 24 port pairs (FP32 blocked solves inside the
    refinement) against the FP64 point."""
    model = plib.Model(*meshgen.box_by_size(0.6, 0.4, 0.3, 0.1))
    eng = plib.Engine(model, retarded=True)
    rng = np.random.default_rng(3)
    ports = [tuple(int(x) for x in rng.choice(model.nv, 2, replace=False)) for _ in range(24)]
    st1, z1, _, _, r1 = eng.solve(4e7, ports, plib.Engine.MIXED, want_solution=False)
    st2, z2, _, _, r2 = eng.solve(4e7, ports, plib.Engine.FP64, want_solution=False)
    eng.close()
    assert st1 == 0 and st2 == 0
    assert r1.backward_error <= 1e-10 and r2.backward_error <= 1e-10
    assert np.max(np.abs(z1 - z2) / np.abs(z2)) < 1e-9


@pytest.mark.parametrize("nports", [8, 20])
def test_engine_backward_error_matches_oracle(require_gpu, nports):
    """The engine's reported backward error (the row-wise residual kernel for
    8 right-hand sides, the MFMA residual for 20) against the oracle's
    harmonic.c:932-969 restatement evaluated on the returned solutions. The
    solutions are the unrefined FP32-factor ones (no refinement, gate lifted),
    so their residuals (~1e-7) stand far above the rounding of the sums and
    the two evaluations must agree closely."""
    xyz, quads = meshgen.box_by_size(0.6, 0.4, 0.3, 0.1)
    model = plib.Model(xyz, quads)
    om = ck.OracleModel(xyz, quads)
    eng = plib.Engine(model, max_refinement_steps=0, backward_error_gate=1.0)
    L, P = eng.static_matrices()
    rng = np.random.default_rng(nports)
    ports = [tuple(int(x) for x in rng.choice(model.nv, 2, replace=False)) for _ in range(nports)]
    f = 3e6
    st, z, I, V, rep = eng.solve(f, ports, plib.Engine.MIXED)
    eng.close()
    assert st == 0 and rep.refinement_steps == 0 and not rep.used_fp64_fallback
    assert rep.backward_error > 1e-12
    Z = om.mna(L, om.pat(P), f)
    etas = []
    for k, (a, b) in enumerate(ports):
        x = np.concatenate([I[k], V[k]])
        rhs = ck.oracle_rhs(model.ne, model.nv, P=P, Is=ck.port_current(model.nv, a, b))
        etas.append(ck.oracle_backward_error(Z, x, rhs, model.ne)[2])
    np.testing.assert_allclose(rep.backward_error, max(etas), rtol=1e-8)


@pytest.mark.parametrize("n", [300, 1030])
def test_real_lu_random(require_gpu, n):
    rng = np.random.default_rng(5)
    A = rng.standard_normal((n, n))
    b = rng.standard_normal(n)
    st, X, _ = plib.real_lu_factor_solve(A, b)
    assert st == 0
    ref = np.linalg.solve(A, b)
    assert np.abs(X[0] - ref).max() <= 1e-10 * np.abs(ref).max()


def test_real_lu_few_rhs_chain(require_gpu):
    """Five right-hand sides through the real chained solve (the variant for
    up to eight per launch), with a partial last 64-row block."""
    rng = np.random.default_rng(15)
    n = 515
    A = rng.standard_normal((n, n))
    B = rng.standard_normal((5, n))
    st, X, _ = plib.real_lu_factor_solve(A, B)
    assert st == 0
    ref = np.linalg.solve(A, B.T).T
    assert np.abs(X - ref).max() <= 1e-10 * np.abs(ref).max()


def _dropin_harmonic(model, frequency, a, b):
    """The reference driver's call sequence (main.c:1349-1465) on the GPU library."""
    lib = plib.load()
    lo = api.InductanceOptions(); lib.inductance_options_default(C.byref(lo), 1)
    po = api.PotentialOptions(); lib.potential_options_default(C.byref(po), 1)
    L, P = api.InductanceMatrix(), api.PotentialMatrix()
    assert lib.inductance_compute_matrix(C.byref(model.mesh), C.byref(model.dual), C.byref(lo), C.byref(L)) == 0
    assert lib.potential_compute_matrix(C.byref(model.mesh), C.byref(model.dual), C.byref(po), C.byref(P)) == 0
    A = api.IncidenceMatrix()
    assert lib.incidence_build(C.byref(model.mesh), C.byref(A)) == 0
    PAT = api.Matrix()
    assert lib.incidence_build_PAT(C.byref(model.mesh), C.byref(P), C.byref(PAT), 1) == 0
    ro = api.ResistanceOptions(); lib.resistance_options_default(C.byref(ro), 1)
    R = api.ResistanceMatrix()
    assert lib.resistance_compute_matrix(C.byref(model.mesh), C.byref(model.dual), C.byref(ro), C.byref(R)) == 0
    system = api.HarmonicSystem()
    lib.harmonic_system_init(C.byref(system))
    assert lib.harmonic_build_system(C.byref(L), C.byref(R), C.byref(A), C.byref(PAT), frequency, 1,
                                     C.byref(system)) == 0
    assert lib.harmonic_factorize(C.byref(system), 1) == 0
    Is = ck.port_current(model.nv, a, b)
    buf = api.complex_buffer(Is)
    ex = api.HarmonicExcitation()
    ex.n_edges, ex.n_nodes = model.ne, model.nv
    ex.node_current = C.cast(buf, C.POINTER(api.Complex))
    sol = api.HarmonicSolution()
    lib.harmonic_solution_init(C.byref(sol))
    assert lib.harmonic_solve(C.byref(system), C.byref(P), C.byref(ex), C.byref(sol), 1) == 0
    res = api.HarmonicResidual()
    assert lib.harmonic_compute_residual(C.byref(system), C.byref(P), C.byref(ex), C.byref(sol), 1,
                                         C.byref(res)) == 0
    I = api.complex_array(sol.edge_current, model.ne)
    V = api.complex_array(sol.node_voltage, model.nv)
    lib.harmonic_solution_free(C.byref(sol))
    lib.harmonic_system_free(C.byref(system))
    for obj, fn in ((L, "inductance_free"), (P, "potential_free"), (A, "incidence_free"), (PAT, "matrix_free"),
                    (R, "resistance_free")):
        getattr(lib, fn)(C.byref(obj))
    return V[a] - V[b], I, V, res.backward_error


@pytest.mark.parametrize("frequency", [1.0, 1e3, 1e6])
def test_dropin_harmonic_plate_msh(require_gpu, frequency):
    """run_tests.py:95-106: backward error <= 1e-9 at 1 Hz / 1 kHz / 1 MHz."""
    with plib.Model(msh=ck.PLATE_MSH) as m:
        zin, I, V, eta = _dropin_harmonic(m, frequency, 0, m.nv - 1)
    print(f"plate.msh f={frequency:g}: Z_in={zin:.9e} eta={eta:.3e}", flush=True)
    assert eta <= 1e-9


def test_dropin_harmonic_plate16_matches_golden(require_gpu):
    g = np.load(os.path.join(GOLD, "plate16.npz"))
    with plib.Model(*meshgen.plate(16, 16)) as m:
        zin, I, V, eta = _dropin_harmonic(m, float(g["frequency"]), 0, 288)
    z_ref = g["z_in"][0]
    print(f"plate16 drop-in: Z_in={zin} ref={z_ref} rel={abs(zin - z_ref) / abs(z_ref):.3e} eta={eta:.3e}")
    assert abs(zin - z_ref) <= 1e-10 * abs(z_ref)
    assert np.abs(I - g["I"]).max() <= 1e-10 * np.abs(g["I"]).max()
    assert np.abs(V - g["V"]).max() <= 1e-10 * np.abs(g["V"]).max()
    assert eta <= 1e-12


@pytest.mark.parametrize("mode", [plib.Engine.FP64, plib.Engine.MIXED])
def test_engine_plate16_matches_golden(require_gpu, mode):
    g = np.load(os.path.join(GOLD, "plate16.npz"))
    with plib.Model(*meshgen.plate(16, 16)) as m:
        eng = plib.Engine(m)
        st, z, I, V, rep = eng.solve(float(g["frequency"]), [(0, 288)], mode)
        eng.close()
    z_ref = g["z_in"][0]
    print(f"engine mode={mode}: Z_in={z[0]} rel={abs(z[0] - z_ref) / abs(z_ref):.3e} eta={rep.backward_error:.3e} "
          f"ir={rep.refinement_steps}", flush=True)
    assert st == 0
    assert abs(z[0] - z_ref) <= 1e-10 * abs(z_ref)
    assert np.abs(I[0] - g["I"]).max() <= 1e-10 * np.abs(g["I"]).max()
    assert np.abs(V[0] - g["V"]).max() <= 1e-10 * np.abs(g["V"]).max()
    assert rep.backward_error <= 1e-10


@pytest.mark.parametrize("frequency", [1e4, 3e6, 2e8])
def test_engine_retarded_box_mixed_vs_fp64_vs_oracle(require_gpu, frequency):
    xyz, quads = meshgen.box(0.6, 0.4, 0.4, 6, 4, 4)
    k = 2 * math.pi * frequency / 299792458.0
    o = ck.OracleModel(xyz, quads)
    try:
        Ls, Ps = o.inductance(), o.potential()
        Lc, Pc = o.retarded(k, 2, Ls, Ps, 4)
        Z = o.mna(None, None, frequency, Lc=Lc, Pc=Pc)
        port = (0, o.nv - 1)
        rhs = ck.oracle_rhs(o.ne, o.nv, Pc=Pc, Is=ck.port_current(o.nv, *port))
        st, fac = ck.oracle_complex_lu(Z)
        x = ck.oracle_complex_solve(fac, rhs)
        z_o = x[o.ne + port[0]] - x[o.ne + port[1]]
    finally:
        o.close()
    with plib.Model(xyz, quads) as m:
        eng = plib.Engine(m, retarded=True)
        st64, z64, I64, _, r64 = eng.solve(frequency, [port], plib.Engine.FP64)
        stm, zm, Im, _, rm = eng.solve(frequency, [port], plib.Engine.MIXED)
        eng.close()
    print(f"f={frequency:g}: oracle {z_o:.10e} fp64 {z64[0]:.10e} mixed {zm[0]:.10e} "
          f"eta64={r64.backward_error:.2e} etam={rm.backward_error:.2e} ir={rm.refinement_steps} "
          f"fallback={rm.used_fp64_fallback}", flush=True)
    assert st64 == 0 and stm == 0
    assert abs(z64[0] - z_o) <= 1e-9 * abs(z_o)
    assert abs(zm[0] - z64[0]) <= 1e-10 * abs(z64[0])
    assert r64.backward_error <= 1e-10 and rm.backward_error <= 1e-10


def test_engine_mixed_sweep_reuses_row_order(require_gpu):
    """Mixed-precision points after the first factor in the row order of the
    previous partial-pivoting factor (no pivot search); every point's
    impedance must still match the FP64 LU to 1e-10 with the backward error
    at or below 1e-10, whether the reused order was accepted or redone."""
    xyz, quads = meshgen.box(1.0, 0.6, 0.6, 10, 6, 6)
    port = (0, len(xyz) - 1)
    freqs = [1e4, 1.1e4, 1e5, 1e6, 1e7, 1e8, 2e8]
    with plib.Model(xyz, quads) as m:
        eng = plib.Engine(m, retarded=True)
        z64 = [eng.solve(f, [port], plib.Engine.FP64)[1][0] for f in freqs]
        eng.close()
        eng = plib.Engine(m, retarded=True)
        reports, zm, widths, rows = [], [], [], []
        for f in freqs:
            st, z, _, _, rep = eng.solve(f, [port], plib.Engine.MIXED)
            assert st == 0
            zm.append(z[0])
            reports.append((rep.pivot_reuse, rep.pivot_refactor, rep.refinement_steps, rep.backward_error,
                            rep.used_fp64_fallback, rep.reuse_backward_error))
            widths.append(rep.block_size)
            rows.append(rep.update_tile_rows)
        eng.close()
    for f, a, b, r in zip(freqs, zm, z64, reports):
        print(f"f={f:g} mixed {a:.12e} fp64 {b:.12e} rel={abs(a - b) / abs(b):.2e} reuse/refactor/ir/eta/fb/eta_reused={r}",
              flush=True)
        assert abs(a - b) <= 1e-10 * abs(b)
        assert r[3] <= 1e-10
    assert reports[0][0] == 0  # the first point has no order to reuse
    assert all(r[0] == 1 for r in reports[1:])
    # the reused order is kept: measured with no refactor over this sweep
    # (from 10 kHz up to 200 MHz in the 10 kHz point's order); one is allowed
    # for a point whose reused-order refinement lands just above its keep
    # threshold
    assert sum(1 for r in reports[1:] if r[1]) <= 1
    # shape autotuner: the first four ordered factors try block widths 128
    # and 64 with split-FP16 update tiles of 64 and 32 rows, the later ones
    # keep one of the four
    print(f"block widths {widths}, update tile rows {rows}", flush=True)
    assert widths[0] == 0 and rows[0] == 0
    assert widths[1:5] == [128, 128, 64, 64] and rows[1:5] == [64, 32, 64, 32]
    assert len(set(zip(widths[5:], rows[5:]))) == 1 and (widths[5], rows[5]) in set(zip(widths[1:5], rows[1:5]))


def test_engine_fixed_block_width(require_gpu):
    """A block width given in the options is used for every ordered factor
    (no autotuning) and gives the same impedances as the autotuned engine."""
    xyz, quads = meshgen.box(0.8, 0.5, 0.5, 8, 5, 5)
    port = (0, len(xyz) - 1)
    freqs = [1e5, 2e5, 3e6, 5e7]
    with plib.Model(xyz, quads) as m:
        out = {}
        for nb in (0, 64):
            eng = plib.Engine(m, retarded=True, block_size=nb)
            res = [eng.solve(f, [port], plib.Engine.MIXED, want_solution=False) for f in freqs]
            eng.close()
            assert all(r[0] == 0 for r in res)
            out[nb] = ([r[1][0] for r in res], [r[4].block_size for r in res])
    assert out[64][1] == [0, 64, 64, 64]
    # the autotuner's first trials: widths 128, 128, 64 (update tiles 64, 32, 64 rows)
    assert out[0][1] == [0, 128, 128, 64]
    np.testing.assert_allclose(out[64][0], out[0][0], rtol=1e-10, atol=0)


@pytest.mark.parametrize("nb", [128, 64])
def test_engine_split_fp16_update_vs_fp32_update(require_gpu, nb):
    """Reused-row-order factors with the split-FP16 MFMA trailing update
    (default) against the FP32 MFMA update and the FP64 LU, over a sweep on a
    box of N = 1538 (12 block columns at 128): every impedance within 1e-10 of
    FP64 with the backward error at or below 1e-10, and refinement taking at
    most four steps more than with the FP32 update over the sweep (step
    counts at the ill-conditioned low-frequency points vary by several steps
    with any rounding change, in either direction)."""
    xyz, quads = meshgen.box(1.0, 0.6, 0.6, 12, 8, 8)
    port = (0, len(xyz) - 1)
    freqs = [1e4, 2e4, 1e5, 1e6, 1e7, 1e8, 2e8]
    res = {}
    with plib.Model(xyz, quads) as m:
        assert m.ne + m.nv > 10 * nb
        for split in (1, 0):
            eng = plib.Engine(m, retarded=True, block_size=nb, split_fp16_update=split)
            res[split] = [eng.solve(f, [port], plib.Engine.MIXED, want_solution=False) for f in freqs]
            eng.close()
        eng = plib.Engine(m, retarded=True)
        z64 = [eng.solve(f, [port], plib.Engine.FP64, want_solution=False)[1][0] for f in freqs]
        eng.close()
    for f, a, b, zr in zip(freqs, res[1], res[0], z64):
        ra, rb = a[4], b[4]
        print(f"f={f:g} split16: steps {ra.refinement_steps} eta {ra.backward_error:.1e} reused {ra.pivot_reuse} "
              f"refactor {ra.pivot_refactor} | fp32: steps {rb.refinement_steps} eta {rb.backward_error:.1e} "
              f"| rel {abs(a[1][0] - zr) / abs(zr):.1e}", flush=True)
        assert a[0] == 0 and b[0] == 0
        assert ra.backward_error <= 1e-10 and not ra.used_fp64_fallback
        assert abs(a[1][0] - zr) <= 1e-10 * abs(zr)
    assert sum(a[4].refinement_steps for a in res[1]) <= sum(b[4].refinement_steps for b in res[0]) + 4


def test_engine_mixed_fp64_fallback_matches_fp64_mode(require_gpu):
    """A mixed point whose refinement cannot reach the gate (no refinement
    steps, gate 1e-13) falls back to the FP64 factor; the mixed path builds
    the equilibrated FP64 matrix only then, from Z0 and the equilibration's
    scales, and must give exactly the FP64 mode's answer (same matrix, same
    factor), on the pivoted first point and on a reused-order point."""
    xyz, quads = meshgen.box(0.8, 0.5, 0.5, 8, 5, 5)
    port = (0, len(xyz) - 1)
    with plib.Model(xyz, quads) as m:
        mixed = plib.Engine(m, retarded=True, max_refinement_steps=0, backward_error_gate=1e-13, gmres_krylov=0)
        fp64 = plib.Engine(m, retarded=True)
        for f in (2e5, 3e7):
            st, z, _, _, rep = mixed.solve(f, [port], plib.Engine.MIXED, want_solution=False)
            st64, z64, _, _, rep64 = fp64.solve(f, [port], plib.Engine.FP64, want_solution=False)
            print(f"f={f:g} fallback={rep.used_fp64_fallback} eta={rep.backward_error:.2e} z={z[0]} z64={z64[0]}",
                  flush=True)
            assert st == 0 and st64 == 0
            assert rep.used_fp64_fallback == 1
            assert z[0] == z64[0]
        mixed.close()
        fp64.close()


def test_engine_from_reference_matrices_via_matrix_market(require_gpu, tmp_path):
    """The reference's own Lp/P of the 16 x 16 plate (golden), written and read
    back as Matrix Market files, drive the engine without assembly: the port
    solution matches the reference's to 1e-10."""
    g = np.load(os.path.join(GOLD, "plate16.npz"))
    ne, nv = int(g["ne"]), int(g["nv"])
    Lg, Pg = np.zeros((ne, ne)), np.zeros((nv, nv))
    Lg[np.triu_indices(ne)] = g["L_upper"]
    Pg[np.triu_indices(nv)] = g["P_upper"]
    Lg = np.triu(Lg) + np.triu(Lg, 1).T
    Pg = np.triu(Pg) + np.triu(Pg, 1).T
    plib.export_dense_mtx(str(tmp_path / "L.mtx"), Lg)
    plib.export_dense_mtx(str(tmp_path / "P.mtx"), Pg)
    L, P = plib.import_dense_mtx(str(tmp_path / "L.mtx")), plib.import_dense_mtx(str(tmp_path / "P.mtx"))
    assert np.array_equal(L, Lg) and np.array_equal(P, Pg)
    with plib.Model(*meshgen.plate(16, 16)) as m:
        eng = plib.Engine(m, static_matrices=(L, P))
        st, z, I, V, rep = eng.solve(float(g["frequency"]), [(0, 288)], plib.Engine.FP64)
        eng.close()
    assert st == 0
    assert abs(z[0] - g["z_in"][0]) <= 1e-10 * abs(g["z_in"][0])
    assert np.abs(I[0] - g["I"]).max() <= 1e-10 * np.abs(g["I"]).max()


@pytest.mark.parametrize("n", [24, 130, 515])
def test_complex_lu_factor_and_pivots_match_oracle(require_gpu, n):
    """The factor, pivots and scales complex_lu_factorize returns to the caller
    (lu->data, lu->pivots, ...) against the oracle, itself bit-identical to the
    reference's complex_lu_factorize: pivots and scales exactly (n = 24 also
    against the reference's own output), the factor entries to 1e-12 of the
    largest (the blocked MFMA update rounds the Schur complements in another
    order than the unblocked elimination)."""
    import tests.golden.make_golden as mg
    g = np.load(os.path.join(GOLD, "residual_lu.npz"))
    A = g["clu_A"] if n == 24 else mg.lu_test_matrix(n, 3 * n)
    st, _, info = plib.complex_lu_factor_solve(A, np.ones(n))
    assert st == 0
    st_o, (LU, piv, rs, cs, ms) = ck.oracle_complex_lu(A)
    assert st_o == 0
    assert np.array_equal(info["pivots"], piv.astype(np.int64))
    assert info["swap_count"] == int(np.sum(piv.astype(np.int64) != np.arange(n)))
    np.testing.assert_allclose(info["row_scale"], rs, rtol=5e-16, atol=0)
    np.testing.assert_allclose(info["column_scale"], cs, rtol=1e-15, atol=0)
    assert abs(info["matrix_scale"] - ms) <= 1e-15 * ms
    err = np.abs(info["data"] - LU).max() / np.abs(LU).max()
    print(f"n={n}: factor max rel deviation {err:.2e}", flush=True)
    assert err <= 1e-12
    if n == 24:
        assert np.array_equal(info["pivots"], g["clu_pivots"])
        assert np.abs(info["data"] - g["clu_factor"]).max() <= 1e-12 * np.abs(g["clu_factor"]).max()


@pytest.mark.parametrize("n", [24, 300])
def test_real_lu_factor_and_pivots_match_oracle(require_gpu, n):
    """lu_factorize's returned factor and pivots against the oracle (and, at
    n = 24, the reference's own output)."""
    import tests.golden.make_golden as mg
    g = np.load(os.path.join(GOLD, "residual_lu.npz"))
    A = g["rlu_A"] if n == 24 else mg.lu_test_matrix(n, 5 * n, complex_=False)
    st, _, (det, LUg, pivg) = plib.real_lu_factor_solve(A, np.ones(n), want_factor=True)
    assert st == 0
    st_o, LU, piv, ms = ck.oracle_real_lu(A)
    assert st_o == 0
    assert np.array_equal(pivg, piv)
    assert np.abs(LUg - LU).max() <= 1e-12 * np.abs(LU).max()
    if n == 24:
        assert np.array_equal(pivg, g["rlu_pivots"])


def test_complex_matrix_vector_multiply_rectangular(require_gpu):
    """complex_matrix.c:542 takes any rows x cols: a 37 x 80 and an 80 x 37
    product against numpy."""
    rng = np.random.default_rng(11)
    for rows, cols in ((37, 80), (80, 37), (1, 5)):
        A = rng.standard_normal((rows, cols)) + 1j * rng.standard_normal((rows, cols))
        x = rng.standard_normal(cols) + 1j * rng.standard_normal(cols)
        y = plib.complex_matvec(A, x)
        ref = A @ x
        assert np.abs(y - ref).max() <= 1e-13 * np.abs(ref).max(), (rows, cols)


@pytest.mark.parametrize("mode", [plib.Engine.FP64, plib.Engine.MIXED])
def test_engine_nonfinite_solution_fails_the_gate(require_gpu, mode):
    """A point whose solution overflows (plane wave of 1e308 V/m at 1 MHz on a
    small box: the oracle's FP64 solution is NaN there) must not pass: the
    NaN residual counts as an infinite backward error and the call returns -3
    (the reference's formula would report 0, NaN dropping out of its maxima)."""
    xyz, quads = meshgen.box(0.6, 0.4, 0.4, 6, 4, 4)
    f, amp = 1.0e6, 1.0e308
    o = ck.OracleModel(xyz, quads)
    try:
        L, P = o.inductance(), o.potential()
        U = o.excitation(ck.oracle_plane_wave(f, 90.0, 0.0, E_amp=amp))
        Z = o.mna(L, o.pat(P), f)
        st, fac = ck.oracle_complex_lu(Z)
        x = ck.oracle_complex_solve(fac, ck.oracle_rhs(o.ne, o.nv, P=P, U=U))
    finally:
        o.close()
    assert not np.isfinite(x).all()  # precondition: the FP64 solution overflows
    with plib.Model(xyz, quads) as m:
        eng = plib.Engine(m)
        st, res, _, rep = eng.monostatic_rcs(f, [90.0], [0.0], mode=mode, E_amp=amp)
        # the same engine still solves a normal point afterwards
        st2, _, _, _, rep2 = eng.solve(f, [(0, m.nv - 1)], mode, want_solution=False)
        eng.close()
    print(f"mode {mode}: status {st} eta {rep.backward_error} fallback {rep.used_fp64_fallback}; then {st2}",
          flush=True)
    assert st == -3 and rep.backward_error == math.inf
    assert st2 == 0 and rep2.backward_error <= 1e-10


def test_engine_rcs_fallback_then_solve_at_same_point(require_gpu):
    """With no refinement steps every mixed point falls back to the FP64
    factor. solve(f) -> monostatic RCS(f) -> solve(f) -> RCS(f) on one engine:
    every call succeeds and gives the FP64 engine's answers (the fallback
    factors the same matrix the same way), so no call reuses a factor that a
    fallback has overwritten."""
    xyz, quads = meshgen.box(0.8, 0.5, 0.5, 8, 5, 5)
    port = (0, len(xyz) - 1)
    f = 2.0e7
    th, ph = [90.0, 60.0, 90.0], [0.0, 45.0, 200.0]
    with plib.Model(xyz, quads) as m:
        fp64 = plib.Engine(m)
        z64 = fp64.solve(f, [port], plib.Engine.FP64, want_solution=False)[1][0]
        s64 = [r.sigma for r in fp64.monostatic_rcs(f, th, ph, mode=plib.Engine.FP64)[1]]
        fp64.close()
        mixed = plib.Engine(m, max_refinement_steps=0, gmres_krylov=0)
        out = []
        for call in ("solve", "rcs", "solve", "rcs"):
            if call == "solve":
                st, z, _, _, rep = mixed.solve(f, [port], plib.Engine.MIXED, want_solution=False)
                out.append((call, st, rep.used_fp64_fallback, rep.factor_cached, z[0]))
            else:
                st, res, _, rep = mixed.monostatic_rcs(f, th, ph, mode=plib.Engine.MIXED)
                out.append((call, st, rep.used_fp64_fallback, rep.factor_cached, [r.sigma for r in res]))
        mixed.close()
    for call, st, fb, cached, val in out:
        print(call, st, fb, cached, flush=True)
        assert st == 0 and fb == 1 and cached == 0
        if call == "solve":
            assert abs(val - z64) <= 1e-12 * abs(z64)
        else:
            np.testing.assert_allclose(val, s64, rtol=1e-12, atol=0)


@pytest.mark.parametrize("kib", [20, 12])
def test_engine_subpanels_of_tall_block_columns(require_gpu, kib):
    """Block columns whose partial-pivoting panel slab does not fit the LDS
    are factored as narrower sub-panels (tall matrices, e.g. N ~ 57000);
    forced here at N = 938 by capping the panel's LDS (20 KiB: 64-column
    sub-panels of the FP64 factor; 12 KiB: 32-column FP64 and 64-column FP32
    sub-panels). The FP64 impedances match the unrestricted factor to 1e-12
    and the mixed point (whose first factor pivots in FP32 sub-panels)
    matches FP64 to 1e-10."""
    xyz, quads = meshgen.box(1.0, 0.6, 0.6, 10, 6, 6)
    port = (0, len(xyz) - 1)
    freqs = [1e6, 5e7]
    with plib.Model(xyz, quads) as m:
        ref = plib.Engine(m, retarded=True)
        z_ref = [ref.solve(f, [port], plib.Engine.FP64, want_solution=False)[1][0] for f in freqs]
        ref.close()
        eng = plib.Engine(m, retarded=True, panel_lds_kib=kib)
        out64 = [eng.solve(f, [port], plib.Engine.FP64, want_solution=False) for f in freqs]
        eng.close()
        eng = plib.Engine(m, retarded=True, panel_lds_kib=kib)
        outm = [eng.solve(f, [port], plib.Engine.MIXED, want_solution=False) for f in freqs]
        eng.close()
    for f, zr, (s64, z64, _, _, r64), (sm, zm, _, _, rm) in zip(freqs, z_ref, out64, outm):
        print(f"f={f:g} cap {kib} KiB: fp64 {abs(z64[0] - zr) / abs(zr):.2e} vs unrestricted, mixed "
              f"{abs(zm[0] - z64[0]) / abs(z64[0]):.2e} vs fp64, eta {r64.backward_error:.1e} / "
              f"{rm.backward_error:.1e}, reuse {rm.pivot_reuse}", flush=True)
        assert s64 == 0 and sm == 0
        assert abs(z64[0] - zr) <= 1e-12 * abs(zr)
        assert abs(zm[0] - z64[0]) <= 1e-10 * abs(z64[0])
        assert r64.backward_error <= 1e-10 and rm.backward_error <= 1e-10


def test_engine_gmres_ir_fallback(require_gpu):
    """When plain refinement stalls above the gate, GMRES-IR on the same FP32
    factor (FP64 residuals; each correction solved by GMRES on the system
    preconditioned with the factor) reaches the gate without the FP64
    refactor. Forced at N = 938 with one plain refinement step and a gate of
    1e-14; with GMRES-IR off the same points refactor in FP64 (whose own
    backward error, ~1e-14 here, need not reach that gate)."""
    xyz, quads = meshgen.box(1.0, 0.6, 0.6, 10, 6, 6)
    port = (0, len(xyz) - 1)
    freqs = [1e6, 3e6, 1e7]
    strict = dict(retarded=True, max_refinement_steps=1, backward_error_gate=1e-14, refinement_target=1e-15)
    with plib.Model(xyz, quads) as m:
        ref = plib.Engine(m, retarded=True)
        z64 = [ref.solve(f, [port], plib.Engine.FP64, want_solution=False)[1][0] for f in freqs]
        ref.close()
        eng = plib.Engine(m, **strict)
        out = [eng.solve(f, [port], plib.Engine.MIXED, want_solution=False) for f in freqs]
        eng.close()
        off = plib.Engine(m, gmres_krylov=0, **strict)
        out_off = [off.solve(f, [port], plib.Engine.MIXED, want_solution=False) for f in freqs]
        off.close()
    for f, zr, (st, z, _, _, r), (st2, _, _, _, r2) in zip(freqs, z64, out, out_off):
        print(f"f={f:g}: steps {r.refinement_steps} gmres {r.gmres_steps}/{r.gmres_iterations} eta "
              f"{r.backward_error:.2e} fallback {r.used_fp64_fallback} | off: fallback {r2.used_fp64_fallback} "
              f"eta {r2.backward_error:.2e} | vs fp64 {abs(z[0] - zr) / abs(zr):.2e}", flush=True)
        assert st == 0 and st2 in (0, -3)
        assert r.used_fp64_fallback == 0 and r.backward_error <= 1e-14
        assert abs(z[0] - zr) <= 1e-10 * abs(zr)
        assert (r.gmres_steps > 0) == (r2.used_fp64_fallback == 1)
    assert any(r.gmres_steps > 0 for _, _, _, _, r in out)
