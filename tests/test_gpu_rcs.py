"""Plane-wave excitation, far field and monostatic RCS on the GPU against the
oracle (itself bit-identical to the compiled reference's
physics/incident_field.c, scattering.c and rcs.c)."""
import math

import numpy as np
import pytest

from peeeeeeeeeeeeeeeeeeec_amd import lib as plib
from peeeeeeeeeeeeeeeeeeec_amd import meshgen
from tests import checker as ck

pytestmark = pytest.mark.gpu

CASES = [(1.0e8, 90.0, 0.0, 1, 2), (3.0e8, 37.0, 211.0, 0, 2), (5.0e7, 120.0, 45.0, 1, 0), (2.0e8, 10.0, 300.0, 0, 1)]


@pytest.fixture(scope="module")
def box():
    xyz, quads = meshgen.box(1.0, 0.6, 0.4, 5, 3, 2)
    o = ck.OracleModel(xyz, quads)
    m = plib.Model(xyz, quads)
    yield o, m
    m.close()
    o.close()


def test_dropin_excitation_matches_oracle(require_gpu, box):
    o, m = box
    for f, th, ph, pol, wt in CASES:
        fg = plib.plane_wave(f, th, ph, pol, 2.5, 0.3, wt)
        fo = ck.oracle_plane_wave(f, th, ph, pol, 2.5, 0.3, wt)
        assert bytes(fg) == bytes(fo)  # host set-up: same arithmetic as the reference
        for integ in (0, 1):
            Ug, Uo = plib.excitation(m, fg, integ), o.excitation(fo, integ)
            assert np.abs(Ug - Uo).max() <= 1e-13 * np.abs(Uo).max(), (f, th, integ)


def test_dropin_far_field_matches_oracle(require_gpu, box):
    o, m = box
    rng = np.random.default_rng(5)
    I = rng.standard_normal(o.ne) + 1j * rng.standard_normal(o.ne)
    for f, th, ph, pol, wt in CASES:
        fo = ck.oracle_plane_wave(f, th, ph, pol, 2.5, 0.3, wt)
        fg = plib.plane_wave(f, th, ph, pol, 2.5, 0.3, wt)
        d = np.array([math.sin(0.7) * math.cos(2.0), math.sin(0.7) * math.sin(2.0), math.cos(0.7)])
        for integ, order in ((0, 6), (1, 6), (1, 3), (1, 16)):
            Fg = plib.radiation_vector(m, I, fo.wave_number, d, integ, order)
            Fo = o.radiation(I, fo.wave_number, d, integ, order)
            assert np.abs(Fg - Fo).max() <= 1e-12 * np.abs(Fo).max(), (f, integ, order)
            rg = plib.rcs_direction(m, fg, I, th + 180.0, ph, integ, order)
            ro = o.rcs_direction(fo, I, th + 180.0, ph, integ, order)
            assert rg.sigma == pytest.approx(ro.sigma, rel=1e-11)
            assert list(rg.direction) == list(ro.direction)


def _oracle_monostatic(o, f, thetas, phis, pol):
    """Reference pipeline per incidence (main_rcs.c) on the retarded MNA:
    excitation, [-U; 0], complex LU, solve, rcs_compute_direction toward
    -k_hat."""
    Ls, Ps = o.inductance(), o.potential()
    k = 2 * math.pi * f / 299792458.0
    Lc, Pc = o.retarded(k, 2, Ls, Ps, 4)
    Z = o.mna(None, None, f, Lc=Lc, Pc=Pc)
    st, fac = ck.oracle_complex_lu(Z)
    assert st == 0
    out = []
    for th, ph in zip(thetas, phis):
        fo = ck.oracle_plane_wave(f, th, ph, pol, 1.0, 0.0, 2)
        U = o.excitation(fo, 1)
        rhs = np.concatenate([-U, np.zeros(o.nv, np.complex128)])
        x = ck.oracle_complex_solve(fac, rhs)
        bt, bp = ck.oracle_backscatter_angles(fo)
        out.append(o.rcs_direction(fo, x[:o.ne], bt, bp, 1, 6))
    return out


@pytest.mark.parametrize("frequency", [3.0e7, 2.0e8])
def test_engine_monostatic_rcs_matches_oracle(require_gpu, frequency):
    """360-degree-style azimuth sweep (12 incidences, theta 90, e_theta) as
    right-hand sides of one factor: FP64 engine vs the oracle pipeline to
    1e-9 in sigma; mixed precision vs FP64 within 1e-9 dB."""
    xyz, quads = meshgen.box(0.6, 0.4, 0.4, 6, 4, 4)
    phis = np.arange(0.0, 360.0, 30.0)
    thetas = np.full_like(phis, 90.0)
    o = ck.OracleModel(xyz, quads)
    try:
        ro = _oracle_monostatic(o, frequency, thetas, phis, 1)
    finally:
        o.close()
    with plib.Model(xyz, quads) as m:
        eng = plib.Engine(m, retarded=True)
        st64, r64, _, rep64 = eng.monostatic_rcs(frequency, thetas, phis, plib.Engine.FP64)
        stm, rm, _, repm = eng.monostatic_rcs(frequency, thetas, phis, plib.Engine.MIXED)
        eng.close()
    assert st64 == 0 and stm == 0
    s_o = np.array([r.sigma for r in ro])
    s64 = np.array([r.sigma for r in r64])
    sm = np.array([r.sigma for r in rm])
    db = np.abs(10 * np.log10(sm / s64))
    print(f"f={frequency:g}: sigma dBsm {np.round(10 * np.log10(s64), 3)}; fp64 vs oracle "
          f"{np.max(np.abs(s64 - s_o) / s_o):.2e}; mixed vs fp64 max {db.max():.2e} dB; "
          f"eta64 {rep64.backward_error:.1e} etam {repm.backward_error:.1e} ir {repm.refinement_steps}", flush=True)
    assert [r.theta_deg for r in r64] == [r.theta_deg for r in ro]
    assert np.max(np.abs(s64 - s_o) / s_o) <= 1e-9
    assert db.max() <= 1e-9
    assert rep64.backward_error <= 1e-10 and repm.backward_error <= 1e-10
