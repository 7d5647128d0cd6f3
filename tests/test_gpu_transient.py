"""Time-domain lightning response by frequency sweep + inverse FFT on the
engine (the C5(a)-style workload at a small size)."""
import math

import numpy as np
import pytest

from peeeeeeeeeeeeeeeeeeec_amd import lib as plib
from peeeeeeeeeeeeeeeeeeec_amd import meshgen
from tests import checker as ck

pytestmark = pytest.mark.gpu


def test_engine_lightning_response(require_gpu):
    xyz, quads = meshgen.plate(8, 8)
    n_freq, df = 64, 4.0e4
    with plib.Model(xyz, quads) as m:
        src = plib.lightning_source(K=30e3, alpha=1.4e4 * 20, beta=6e6, strike=0, ret=m.nv - 1)
        eng = plib.Engine(m)
        st, t, i, v, z, reps = eng.lightning_response(src, n_freq, df, plib.Engine.FP64)
        assert st == 0
        # the sweep is the engine's own port solve at each frequency
        for k in (0, 17, n_freq - 1):
            st1, z1, *_ = eng.solve((k + 1) * df, [(0, m.nv - 1)], plib.Engine.FP64, want_solution=False)
            assert st1 == 0 and z1[0] == z[k]
        stm, tm, im, vm, zm, repm = eng.lightning_response(src, n_freq, df, plib.Engine.MIXED)
        eng.close()
    assert stm == 0
    assert np.abs(zm - z).max() <= 1e-10 * np.abs(z).max()
    assert np.abs(vm - v).max() <= 1e-9 * np.abs(v).max()
    assert max(r.backward_error for r in reps + repm) <= 1e-10
    # superposition: the response is the host inverse transform of Z_in I
    t2, i2, v2 = plib.transient_from_spectrum(src, df, z)
    assert np.array_equal(v, v2) and np.array_equal(t, t2)
    # Z_in against the oracle at three frequencies of the sweep
    o = ck.OracleModel(xyz, quads)
    try:
        Ls, Ps = o.inductance(), o.potential()
        PAT = o.pat(Ps)
        for k in (0, 31, 63):
            f = (k + 1) * df
            Z = o.mna(Ls, PAT, f)
            st_o, fac = ck.oracle_complex_lu(Z)
            rhs = ck.oracle_rhs(o.ne, o.nv, P=Ps, Is=ck.port_current(o.nv, 0, o.nv - 1))
            x = ck.oracle_complex_solve(fac, rhs)
            zo = x[o.ne] - x[o.ne + o.nv - 1]
            assert abs(z[k] - zo) <= 1e-10 * abs(zo)
    finally:
        o.close()
    print(f"lightning response: peak |v| {np.abs(v).max():.3e} V at t={t[np.argmax(np.abs(v))]:.3e} s, "
          f"peak i {i.max():.3e} A", flush=True)
