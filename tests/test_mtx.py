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
