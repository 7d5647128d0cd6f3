"""Lightning source (drop-in for physics/lightning.c) and the frequency-domain
time response of a port (sweep + inverse FFT), on the CPU: the host pieces
against the compiled reference and against closed-form circuit responses."""
import ctypes as C
import math

import numpy as np
import pytest

from peeeeeeeeeeeeeeeeeeec_amd import ctypes_api as api
from peeeeeeeeeeeeeeeeeeec_amd import lib as plib
from peeeeeeeeeeeeeeeeeeec_amd import meshgen
from tests import checker as ck


def test_lightning_struct_layout():
    assert C.sizeof(api.LightningSource) == 56  # gcc on the reference's lightning.h


@pytest.mark.skipif(not ck.ref_available(), reason="oracle/_ref not built")
def test_lightning_functions_bit_exact_against_reference_build():
    ref, lib = ck.ref(), plib.load()
    S = C.POINTER(api.LightningSource)
    for name, res, args in (("lightning_source_default", None, [S]), ("lightning_current", C.c_double, [S, C.c_double]),
                            ("lightning_peak", C.c_int, [S, ck._dp, ck._dp]),
                            ("lightning_source_validate", C.c_int, [S, C.c_size_t]),
                            ("lightning_build_node_current", C.c_int, [S, C.c_size_t, C.c_double, ck._dp]),
                            ("lightning_find_nearest_node", C.c_int, [ck._M, ck._dp, C.POINTER(C.c_size_t), ck._dp])):
        getattr(ref, name).restype = res
        getattr(ref, name).argtypes = args
    a, b = api.LightningSource(), api.LightningSource()
    ref.lightning_source_default(C.byref(a))
    lib.lightning_source_default(C.byref(b))
    assert bytes(a) == bytes(b)
    for s in (a, b):
        s.K, s.alpha, s.beta, s.delay, s.strike_node, s.return_node = 30e3, 1.4e4, 6e6, 2e-6, 3, 17
    for t in (0.0, 1e-6, 2e-6, 2.5e-6, 1e-5, 3e-4, 1e-2):
        assert ref.lightning_current(C.byref(a), t) == lib.lightning_current(C.byref(b), t)
    tp = [C.c_double(), C.c_double()]
    tq = [C.c_double(), C.c_double()]
    assert ref.lightning_peak(C.byref(a), C.byref(tp[0]), C.byref(tp[1])) == 0
    assert lib.lightning_peak(C.byref(b), C.byref(tq[0]), C.byref(tq[1])) == 0
    assert (tp[0].value, tp[1].value) == (tq[0].value, tq[1].value)
    x, y = np.zeros(20), np.ones(20)
    assert ref.lightning_build_node_current(C.byref(a), 20, 3e-6, x.ctypes.data_as(ck._dp)) == 0
    assert lib.lightning_build_node_current(C.byref(b), 20, 3e-6, y.ctypes.data_as(ck._dp)) == 0
    assert np.array_equal(x, y)
    for bad in (dict(K=-1.0), dict(beta=1e3), dict(delay=-1.0), dict(strike_node=25), dict(return_node=3)):
        for s in (a, b):
            s.K, s.alpha, s.beta, s.delay, s.strike_node, s.return_node = 30e3, 1.4e4, 6e6, 0.0, 3, 17
            for k, v in bad.items():
                setattr(s, k, v)
        assert ref.lightning_source_validate(C.byref(a), 20) == lib.lightning_source_validate(C.byref(b), 20) == -1
    xyz, quads = meshgen.plate(5, 4)
    with plib.Model(xyz, quads) as m:
        p = np.array([0.13, -0.21, 0.05])
        na, nb, da, db = C.c_size_t(), C.c_size_t(), C.c_double(), C.c_double()
        assert ref.lightning_find_nearest_node(C.byref(m.mesh), p.ctypes.data_as(ck._dp), C.byref(na),
                                               C.byref(da)) == 0
        assert lib.lightning_find_nearest_node(C.byref(m.mesh), p.ctypes.data_as(ck._dp), C.byref(nb),
                                               C.byref(db)) == 0
        assert (na.value, da.value) == (nb.value, db.value)


def test_lightning_spectrum_is_the_fourier_transform():
    """I(f) against a direct quadrature of int i(t) e^{-j 2 pi f t} dt."""
    s = plib.lightning_source(K=1.0, alpha=1e4, beta=2e5, delay=1e-5)
    t = np.linspace(0.0, 2e-3, 400001)
    i = np.where(t >= 1e-5, np.exp(-1e4 * (t - 1e-5)) - np.exp(-2e5 * (t - 1e-5)), 0.0)
    for f in (0.0, 300.0, 5e3, 4e4):
        I = api.Complex()
        assert plib.load().peec_lightning_spectrum(C.byref(s), f, C.byref(I)) == 0
        g = i * np.exp(-2j * math.pi * f * t)
        q = np.trapezoid(g, t) if hasattr(np, "trapezoid") else np.trapz(g, t)
        assert abs(complex(I.re, I.im) - q) <= 2e-4 * abs(q) + 1e-12


@pytest.mark.parametrize("R,Lh", [(2.0, 0.0), (0.5, 3e-5), (0.0, 1e-4)])
def test_transient_from_spectrum_matches_series_rl(R, Lh):
    """Z(f) = R + j 2 pi f L gives v(t) = R i(t) + L di/dt in closed form;
    the sweep-and-inverse-FFT response reproduces it away from the window
    edges."""
    s = plib.lightning_source(K=30e3, alpha=1e4, beta=1e5, strike=0, ret=1)
    n, df = 4096, 100.0
    f = df * np.arange(1, n + 1)
    z = R + 2j * math.pi * f * Lh
    t, i, v = plib.transient_from_spectrum(s, df, z)
    exact = R * i + Lh * 30e3 * (-1e4 * np.exp(-1e4 * t) + 1e5 * np.exp(-1e5 * t))
    sel = (t > 2e-4) & (t < 5e-3)
    scale = np.abs(exact).max()
    assert np.abs(v[sel] - exact[sel]).max() <= 2e-3 * scale, np.abs(v[sel] - exact[sel]).max() / scale
    assert t[1] == pytest.approx(1.0 / (2 * n * df))
    assert np.array_equal(i, [plib.load().lightning_current(C.byref(s), x) for x in t])


def test_transient_rejects_bad_sizes():
    s = plib.lightning_source(K=1.0, alpha=1e4, beta=1e5, strike=0, ret=1)
    with pytest.raises(RuntimeError):
        plib.transient_from_spectrum(s, 100.0, np.ones(1000, complex))  # not a power of two
