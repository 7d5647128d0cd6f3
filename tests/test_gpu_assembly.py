"""GPU assembly parity (FP64) against the golden vectors of the reference
and the oracle: static Lp/P to 1e-12 relative per entry, retarded matrices
to 1e-11 of the largest entry."""
from __future__ import annotations

import os

import numpy as np
import pytest

from peeeeeeeeeeeeeeeeeeec_amd import ctypes_api as api
from peeeeeeeeeeeeeeeeeeec_amd import lib as plib
from peeeeeeeeeeeeeeeeeeec_amd import meshgen
from tests import checker as ck

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
ENTRY_RTOL = 1e-12


def rel_err(a, b):
    """Entry-wise relative error, zeros compared absolutely to the scale."""
    scale = np.abs(b).max()
    denom = np.maximum(np.abs(b), 1e-300)
    err = np.where(b != 0, np.abs(a - b) / denom, np.abs(a - b) / scale)
    return err


def report(name, a, b):
    e = rel_err(a, b)
    print(f"{name}: max entry rel err {e.max():.3e}, entries > 1e-12: {(e > ENTRY_RTOL).sum()} of {e.size}",
          flush=True)
    return e


def test_plate_msh_static_matches_golden(require_gpu):
    g = np.load(os.path.join(GOLD, "plate_msh.npz"))
    with plib.Model(msh=ck.PLATE_MSH) as m:
        L = m.inductance()
        P = m.potential()
    assert report("L plate.msh", L, g["L"]).max() <= ENTRY_RTOL
    assert report("P plate.msh", P, g["P"]).max() <= ENTRY_RTOL
    assert np.array_equal(L[g["L"] == 0], g["L"][g["L"] == 0])


def test_plate16_static_matches_golden(require_gpu):
    g = np.load(os.path.join(GOLD, "plate16.npz"))
    with plib.Model(*meshgen.plate(16, 16)) as m:
        L = m.inductance()
        P = m.potential()
    iu, ju = np.triu_indices(m.ne), np.triu_indices(m.nv)
    assert report("L plate16", L[iu], g["L_upper"]).max() <= ENTRY_RTOL
    assert report("P plate16", P[ju], g["P_upper"]).max() <= ENTRY_RTOL
    assert np.array_equal(L, L.T) and np.array_equal(P, P.T)


def test_box_static_matches_oracle(require_gpu):
    xyz, quads = meshgen.box(1.0, 0.6, 0.4, 5, 3, 2)
    o = ck.OracleModel(xyz, quads)
    try:
        Lo, Po = o.inductance(), o.potential()
    finally:
        o.close()
    with plib.Model(xyz, quads) as m:
        L, P = m.inductance(), m.potential()
    assert report("L box", L, Lo).max() <= ENTRY_RTOL
    assert report("P box", P, Po).max() <= ENTRY_RTOL


def test_piece_pairs_match_golden(require_gpu):
    import ctypes as C
    g = np.load(os.path.join(GOLD, "pairs.npz"))
    lib = plib.load()
    for a, b, val in zip(g["a"], g["b"], g["value"]):
        pa, pb = api.QuadPatch(), api.QuadPatch()
        for i in range(12):
            pa.points[i] = a[i]
            pb.points[i] = b[i]
        same = np.array_equal(a, b)
        v = lib.quad_patch_kernel_1_over_r(C.byref(pa), C.byref(pa) if same else C.byref(pb), 4, 10, 12, 8, 0.5)
        assert abs(v - val) <= ENTRY_RTOL * abs(val), (v, val)
    bad = lib.quad_patch_kernel_1_over_r(C.byref(pa), C.byref(pb), 7, 10, 12, 8, 0.5)
    assert np.isnan(bad)


def test_compute_pair_matches_matrix(require_gpu):
    import ctypes as C
    g = np.load(os.path.join(GOLD, "plate_msh.npz"))
    lib = plib.load()
    with plib.Model(msh=ck.PLATE_MSH) as m:
        lo = api.InductanceOptions()
        lib.inductance_options_default(C.byref(lo), 1)
        po = api.PotentialOptions()
        lib.potential_options_default(C.byref(po), 1)
        for a, b in ((0, 0), (0, 1), (3, 40), (17, 17), (5, 179)):
            v = C.c_double()
            assert lib.inductance_compute_pair(C.byref(m.mesh), C.byref(m.dual), a, b, C.byref(lo), C.byref(v)) == 0
            assert abs(v.value - g["L"][a, b]) <= ENTRY_RTOL * abs(g["L"][a, b]) + 0.0
        for i, j in ((0, 0), (0, 1), (10, 55), (99, 99)):
            v = C.c_double()
            assert lib.potential_compute_pair(C.byref(m.mesh), C.byref(m.dual), i, j, C.byref(po), C.byref(v)) == 0
            assert abs(v.value - g["P"][i, j]) <= ENTRY_RTOL * abs(g["P"][i, j])


@pytest.mark.parametrize("k", [0.0, 0.7, 4.19])
def test_retarded_matches_oracle(require_gpu, k):
    xyz, quads = meshgen.box(0.6, 0.4, 0.4, 3, 2, 2)
    o = ck.OracleModel(xyz, quads)
    try:
        Ls, Ps = o.inductance(), o.potential()
        Lo, Po = o.retarded(k, 2, Ls, Ps, 4)
    finally:
        o.close()
    with plib.Model(xyz, quads) as m:
        L, P = m.retarded(k, 2, 4)
    el = np.abs(L - Lo).max() / np.abs(Lo).max()
    ep = np.abs(P - Po).max() / np.abs(Po).max()
    print(f"retarded k={k}: L {el:.3e}, P {ep:.3e}", flush=True)
    assert el <= 1e-11 and ep <= 1e-11


@pytest.mark.parametrize("k", [0.0, 1e-4, 0.05, 1.0, 3.0, 6.0])
def test_retarded_from_moments_matches_quadrature(require_gpu, k):
    """Retarded Lp(w)/P(w) from the stored power-series moments of the smooth
    remainder against the quadrature pass of the same engine options and the
    oracle. The 0.6 x 0.4 x 0.4 m box has a diameter of 0.825 m, so with 26
    stored terms k <= 3 takes the moments (k * diameter <= 2.56) and k = 6 the
    quadrature. Real and imaginary parts are checked separately: the
    imaginary part (radiation loss) is tiny next to the static values at low k."""
    xyz, quads = meshgen.box(0.6, 0.4, 0.4, 6, 4, 4)
    with plib.Model(xyz, quads) as m:
        out = {}
        for terms in (26, 0):
            eng = plib.Engine(m, retarded=True, smooth_moments=terms, shifted_moments=0)
            out[terms] = eng.retarded_matrices(k)
            eng.close()
    o = ck.OracleModel(xyz, quads)
    try:
        Lo, Po = o.retarded(k, 2, o.inductance(), o.potential(), 4)
    finally:
        o.close()
    for name, a, b, c in (("L", out[26][0], out[0][0], Lo), ("P", out[26][1], out[0][1], Po)):
        assert np.array_equal(a, a.T) and np.array_equal(b, b.T)
        e_re = np.abs(a.real - b.real).max() / np.abs(b.real).max()
        im_scale = np.abs(b.imag).max()
        e_im = np.abs(a.imag - b.imag).max() / im_scale if im_scale > 0 else np.abs(a.imag).max()
        e_or = np.abs(a - c).max() / np.abs(c).max()
        print(f"k={k} {name}: moments vs quadrature re {e_re:.2e} im {e_im:.2e}; vs oracle {e_or:.2e}", flush=True)
        assert e_re <= 1e-14 and e_im <= 1e-12 and e_or <= 1e-11
        if k == 0.0:
            assert np.array_equal(a.imag, np.zeros_like(a.imag))


def test_engine_points_from_moments(require_gpu):
    """Sweep points on the moment path (report.moment_terms > 0) and above its
    range (quadrature) give the impedances of an engine without moments."""
    xyz, quads = meshgen.box(0.8, 0.5, 0.5, 8, 5, 5)
    port = (0, len(xyz) - 1)
    freqs = [1e4, 3e6, 1e8, 2e8, 6e8]
    res = {}
    with plib.Model(xyz, quads) as m:
        for terms in (-1, 0):
            eng = plib.Engine(m, retarded=True, smooth_moments=terms, shifted_moments=0)
            res[terms] = [eng.solve(f, [port], plib.Engine.MIXED, want_solution=False) for f in freqs]
            eng.close()
    used = [r[4].moment_terms for r in res[-1]]
    print("moment terms per point:", used, flush=True)
    assert all(r[0] == 0 for r in res[-1] + res[0])
    assert all(r[4].moment_terms == 0 for r in res[0])
    # diameter 1.06 m: k * diameter = 2.2 at 100 MHz (moments), 13 at 600 MHz (quadrature)
    assert used[0] >= 1 and used[1] > used[0] and used[2] > used[1] and used[-1] == 0
    for a, b in zip(res[-1], res[0]):
        assert abs(a[1][0] - b[1][0]) <= 1e-10 * abs(b[1][0])
        assert a[4].backward_error <= 1e-10


def test_retarded_from_shifted_moments_matches_quadrature(require_gpu):
    """Points above the range of the moments at 0 (k * diameter > 2.56 on the
    0.825 m box) from complex moments around a shifted wave number k0 against
    the quadrature pass and the oracle: real and imaginary parts to 1e-14 /
    1e-12 as for the moments at 0. The first point builds the center (k0 above
    k), the next two reuse it, a point below its window rebuilds it."""
    xyz, quads = meshgen.box(0.6, 0.4, 0.4, 6, 4, 4)
    ks = [6.0, 8.0, 11.0, 4.0]
    with plib.Model(xyz, quads) as m:
        sh = plib.Engine(m, retarded=True, smooth_moments=26, shifted_moments=30)
        qd = plib.Engine(m, retarded=True, smooth_moments=0, shifted_moments=0)
        res = [(k, sh.retarded_matrices(k), qd.retarded_matrices(k)) for k in ks]
        sh.close()
        qd.close()
    o = ck.OracleModel(xyz, quads)
    try:
        Ls, Ps = o.inductance(), o.potential()
        for k, a, b in res:
            Lo, Po = o.retarded(k, 2, Ls, Ps, 4)
            for name, x, y, c in (("L", a[0], b[0], Lo), ("P", a[1], b[1], Po)):
                e_re = np.abs(x.real - y.real).max() / np.abs(y.real).max()
                e_im = np.abs(x.imag - y.imag).max() / np.abs(y.imag).max()
                e_or = np.abs(x - c).max() / np.abs(c).max()
                print(f"k={k} {name}: shifted vs quadrature re {e_re:.2e} im {e_im:.2e}; vs oracle {e_or:.2e}",
                      flush=True)
                assert np.array_equal(x, x.T)
                assert e_re <= 1e-14 and e_im <= 1e-12 and e_or <= 1e-11
    finally:
        o.close()


def test_retarded_from_many_moments_matches_quadrature(require_gpu):
    """56 stored terms on the 0.825 m box cover k * diameter up to ~10.5: at
    k = 11.5 (k * diameter = 9.5) the alternating series loses up to
    e^9.5 ~ 1.3e4 to cancellation, so the moment path matches the quadrature
    pass to ~1e-12 (5e-12 allowed, against the largest real / imaginary part)
    and the oracle to 1e-11; at k = 3 as closely as with 26 terms."""
    xyz, quads = meshgen.box(0.6, 0.4, 0.4, 6, 4, 4)
    ks = [3.0, 11.5]
    with plib.Model(xyz, quads) as m:
        mo = plib.Engine(m, retarded=True, smooth_moments=56, shifted_moments=0)
        qd = plib.Engine(m, retarded=True, smooth_moments=0, shifted_moments=0)
        res = [(k, mo.retarded_matrices(k), qd.retarded_matrices(k)) for k in ks]
        mo.close()
        qd.close()
    o = ck.OracleModel(xyz, quads)
    try:
        Ls, Ps = o.inductance(), o.potential()
        for k, a, b in res:
            Lo, Po = o.retarded(k, 2, Ls, Ps, 4)
            tol = 1e-14 if k * 0.825 <= 2.56 else 5e-12
            for name, x, y, c in (("L", a[0], b[0], Lo), ("P", a[1], b[1], Po)):
                e_re = np.abs(x.real - y.real).max() / np.abs(y.real).max()
                e_im = np.abs(x.imag - y.imag).max() / np.abs(y.imag).max()
                e_or = np.abs(x - c).max() / np.abs(c).max()
                print(f"k={k} {name}: 56 moments vs quadrature re {e_re:.2e} im {e_im:.2e}; vs oracle {e_or:.2e}",
                      flush=True)
                assert np.array_equal(x, x.T)
                assert e_re <= tol and e_im <= max(tol, 1e-12) and e_or <= 1e-11
    finally:
        o.close()


def test_engine_points_from_shifted_moments(require_gpu):
    """A sweep above the range of the moments at 0: the points report the
    shifted center they used and give the impedances of an engine without
    moments."""
    xyz, quads = meshgen.box(0.8, 0.5, 0.5, 8, 5, 5)
    port = (0, len(xyz) - 1)
    freqs = [2.5e8, 3e8, 4e8]
    res = {}
    with plib.Model(xyz, quads) as m:
        for on in (1, 0):
            eng = plib.Engine(m, retarded=True, smooth_moments=26 if on else 0, shifted_moments=-on)
            res[on] = [eng.solve(f, [port], plib.Engine.MIXED, want_solution=False) for f in freqs]
            eng.close()
    centers = [r[4].moment_center for r in res[1]]
    print("centers:", centers, "terms:", [r[4].moment_terms for r in res[1]], flush=True)
    assert all(c > 0 for c in centers) and centers[0] == centers[1]
    assert all(r[4].moment_center == 0 and r[4].moment_terms == 0 for r in res[0])
    for a, b in zip(res[1], res[0]):
        assert a[0] == 0 and b[0] == 0
        assert abs(a[1][0] - b[1][0]) <= 1e-10 * abs(b[1][0])


def test_gpu_matrices_export_to_exact_capacitance_network(require_gpu, tmp_path):
    """The reference's SPICE export check (tests/run_tests.py:166-239) on the
    GPU-assembled Lp/P of meshes/plate.msh: written by matrix_export_peec_set
    and read back by scipy.io.mmread they are the assembled numbers bit for
    bit, and the capacitance network built from P reconstructs P^-1 to
    1e-10."""
    from tests.test_mtx import _capacitance_network_error, _export_set

    with plib.Model(msh=ck.PLATE_MSH) as m:
        L, P = m.inductance(), m.potential()
        read, A = _export_set(tmp_path, m, L, P)
    assert np.array_equal(read["L"], L) and np.array_equal(read["P"], P) and np.array_equal(read["A"], A)
    err = _capacitance_network_error(read["P"])
    print(f"capacitance network reconstruction error {err:.3e}", flush=True)
    assert np.isfinite(err) and err <= 1e-10
