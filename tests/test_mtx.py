"""Matrix Market interchange (drop-in for io/matrix_export.c): files are
byte-identical to the reference build's; the import reads them back
exactly."""
import ctypes as C
import os

import numpy as np
import pytest

from peeeeeeeeeeeeeeeeeeec_amd import ctypes_api as api
from peeeeeeeeeeeeeeeeeeec_amd import lib as plib
from peeeeeeeeeeeeeeeeeeec_amd import meshgen
from tests import checker as ck


def _ref_sigs(ref):
    ref.matrix_export_dense_mtx.argtypes = [C.c_char_p, ck._dp, C.c_size_t, C.c_size_t]
    ref.matrix_export_dense_mtx.restype = C.c_int
    ref.matrix_export_peec_set.argtypes = [C.c_char_p, ck._M, C.POINTER(api.InductanceMatrix),
                                           C.POINTER(api.PotentialMatrix), C.POINTER(api.Matrix),
                                           C.POINTER(api.ResistanceMatrix)]
    ref.matrix_export_peec_set.restype = C.c_int


@pytest.mark.skipif(not ck.ref_available(), reason="oracle/_ref not built")
def test_exported_files_are_byte_identical(tmp_path):
    ref, lib = ck.ref(), plib.load()
    _ref_sigs(ref)
    rng = np.random.default_rng(2)
    M = rng.standard_normal((37, 23)) * 10.0 ** rng.integers(-30, 30, size=(37, 23))
    M[3, 4], M[5, 6] = 0.0, -0.0
    a, b = str(tmp_path / "a.mtx"), str(tmp_path / "b.mtx")
    assert ref.matrix_export_dense_mtx(a.encode(), M.ctypes.data_as(ck._dp), 37, 23) == 0
    plib.export_dense_mtx(b, M)
    assert open(a, "rb").read() == open(b, "rb").read()
    back = plib.import_dense_mtx(b)
    assert np.array_equal(back, M) and np.array_equal(np.signbit(back), np.signbit(M))
    # the full set on a small model (L, A, P, PAT, R)
    xyz, quads = meshgen.plate(3, 2)
    o = ck.OracleModel(xyz, quads)
    try:
        Ln, Pn = o.inductance(), o.potential()
        PATn = o.pat(Pn)
        L = api.InductanceMatrix(n=o.ne, data=api.c_malloc_copy(Ln, C.c_double))
        P = api.PotentialMatrix(n=o.nv, data=api.c_malloc_copy(Pn, C.c_double))
        PAT = api.Matrix(rows=o.nv, cols=o.ne, data=api.c_malloc_copy(PATn, C.c_double))
        R = api.ResistanceMatrix(n=o.ne, diagonal=api.c_malloc_copy(np.zeros(o.ne), C.c_double))
        da, db = str(tmp_path / "ref" / "set"), str(tmp_path / "ours" / "set")
        assert ref.matrix_export_peec_set(da.encode(), C.byref(o.mesh), C.byref(L), C.byref(P), C.byref(PAT),
                                          C.byref(R)) == 0
        assert lib.matrix_export_peec_set(db.encode(), C.byref(o.mesh), C.byref(L), C.byref(P), C.byref(PAT),
                                          C.byref(R)) == 0
        for name in ("L.mtx", "A.mtx", "P.mtx", "PAT.mtx", "R.mtx"):
            assert open(os.path.join(da, name), "rb").read() == open(os.path.join(db, name), "rb").read(), name
        assert np.array_equal(plib.import_dense_mtx(os.path.join(db, "L.mtx")), Ln)
    finally:
        o.close()


def test_import_rejects_other_formats(tmp_path):
    p = tmp_path / "x.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 1.0\n")
    with pytest.raises(RuntimeError):
        plib.import_dense_mtx(str(p))
    p.write_text("%%MatrixMarket matrix array real general\n2 2\n1.0\n2.0\n3.0\n")
    with pytest.raises(RuntimeError):
        plib.import_dense_mtx(str(p))


def _export_set(tmp_path, model, L, P):
    """matrix_export_peec_set of (L, A, P, PAT, R = 0) for `model`, read back
    with scipy's Matrix Market reader."""
    from scipy.io import mmread

    lib = plib.load()
    ne, nv = L.shape[0], P.shape[0]
    A = np.zeros((ne, nv))
    e = model.edges()
    A[np.arange(ne), e[:, 0]], A[np.arange(ne), e[:, 1]] = -1.0, 1.0
    PAT = P @ A.T
    Lm = api.InductanceMatrix(n=ne, data=api.c_malloc_copy(np.ascontiguousarray(L), C.c_double))
    Pm = api.PotentialMatrix(n=nv, data=api.c_malloc_copy(np.ascontiguousarray(P), C.c_double))
    PATm = api.Matrix(rows=nv, cols=ne, data=api.c_malloc_copy(np.ascontiguousarray(PAT), C.c_double))
    Rm = api.ResistanceMatrix(n=ne, diagonal=api.c_malloc_copy(np.zeros(ne), C.c_double))
    d = str(tmp_path / "set")
    assert lib.matrix_export_peec_set(d.encode(), C.byref(model.mesh), C.byref(Lm), C.byref(Pm), C.byref(PATm),
                                      C.byref(Rm)) == 0
    read = {}
    for name in ("L", "P", "A", "R"):
        v = mmread(os.path.join(d, name + ".mtx"))
        read[name] = np.asarray(v.toarray() if hasattr(v, "toarray") else v, dtype=float)
    return read, A


def _capacitance_network_error(P):
    """The reference's SPICE check (tests/run_tests.py:166-239): the nodal
    capacitance matrix C = P^-1 written as a network of two-terminal
    capacitors (c_ij = -C_ij between nodes i and j, sum_j C_ij from node i to
    ground) is stamped back into a nodal matrix; returns the relative
    infinity-norm error of that reconstruction against P^-1."""
    Cm = np.linalg.solve(P, np.eye(P.shape[0]))
    n = P.shape[0]
    rec = np.zeros_like(Cm)
    for i in range(n):
        g = Cm[i].sum()  # capacitor from node i to ground
        rec[i, i] += g
        for j in range(i + 1, n):
            c = -Cm[i, j]
            rec[i, i] += c
            rec[j, j] += c
            rec[i, j] -= c
            rec[j, i] -= c
    return np.linalg.norm(rec - Cm, np.inf) / np.linalg.norm(Cm, np.inf)


def test_exported_reference_matrices_give_exact_capacitance_network(tmp_path):
    """The reference's own Lp/P of meshes/plate.msh (golden), exported by this
    library and read back by scipy.io.mmread, are the same numbers bit for
    bit, A holds one -1/+1 pair per edge, and the capacitance network built
    from the exported P reconstructs P^-1 to 1e-10 (run_tests.py:209-226)."""
    g = np.load(os.path.join(ck.ROOT, "tests", "golden", "plate_msh.npz"))
    with plib.Model(msh=ck.PLATE_MSH) as m:
        read, A = _export_set(tmp_path, m, g["L"], g["P"])
    assert np.array_equal(read["L"], g["L"]) and np.array_equal(read["P"], g["P"])
    assert np.array_equal(read["A"], A) and not read["R"].any()
    assert np.all(np.abs(read["A"]).sum(axis=1) == 2) and np.all(read["A"].sum(axis=1) == 0)
    err = _capacitance_network_error(read["P"])
    print(f"capacitance network reconstruction error {err:.3e}")
    assert np.isfinite(err) and err <= 1e-10
