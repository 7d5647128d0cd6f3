"""The theta-scheme transient drop-in (transient.h) on the GPU: the reference's
known answers (tests/test_core.c:68-177 of the reference), the system matrix
bit-identical to the oracle's (itself bit-identical to the compiled
reference), stepped solutions against the oracle, and the incidence /
potential operators the transient uses."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from peeeeeeeeeeeeeeeeeeec_amd import ctypes_api as api
from peeeeeeeeeeeeeeeeeeec_amd import lib as plib
from peeeeeeeeeeeeeeeeeeec_amd import meshgen
from tests import checker as ck
from tests.test_transient_theta import BE, TR, drive, incidence_dense, shunt_energy, time_order

pytestmark = pytest.mark.gpu


def gpu_make(L, R, A, P, scheme, dt, shunt=None):
    return ck.TransientRun(plib.load(), L, R, A, P, scheme, dt, shunt)


def mesh_operators(xyz, quads, seed=7):
    o = ck.OracleModel(xyz, quads)
    try:
        L, P = o.inductance(), o.potential()
        A = incidence_dense(o.edges(), o.nv)
    finally:
        o.close()
    R = np.random.default_rng(seed).uniform(1e-3, 1e-2, L.shape[0])
    return L, R, A, P


@pytest.mark.parametrize("scheme", [BE, TR])
def test_oscillator_time_order_and_energy(require_gpu, scheme):
    """test_core.c:79-130 through the GPU drop-in: error ratios 2 (backward
    Euler) and 4 (trapezoidal) per halving of dt, trapezoidal energy conserved
    and backward-Euler energy non-increasing to 1e-11, total charge zero."""
    time_order(gpu_make, scheme)


def test_shunt_energy_balance(require_gpu):
    """test_core.c:133-170 through the GPU drop-in (stage 2)."""
    shunt_energy(gpu_make)


@pytest.mark.parametrize("scheme,stage2", [(BE, False), (TR, False), (TR, True)])
def test_box_transient_against_oracle(require_gpu, scheme, stage2):
    """A 1 x 0.5 x 0.5 m box: the GPU-filled system matrix equals the
    oracle's bit for bit (same expressions per entry); eight steps with
    changing node sources and edge voltages: right-hand sides within 1e-13
    and solutions within 1e-11 of the largest entry (the GPU sums the
    products in another order and factors with the blocked LU)."""
    L, R, A, P = mesh_operators(*meshgen.box(1.0, 0.5, 0.5, 4, 2, 2))
    ne, nv = A.shape
    dt = 2e-9
    shunt = (0, nv - 1, 2.0) if stage2 else None
    gpu = gpu_make(L, R, A, P, scheme, dt, shunt)
    ora = ck.OracleTransient(L, R, A, P, scheme, dt, shunt)
    try:
        assert np.array_equal(gpu.matrix(), ora.M)
        a = drive(gpu, 8, ne, nv, 3, stage2)
        b = drive(ora, 8, ne, nv, 3, stage2)
    finally:
        gpu.close()
    for (Ig, Vg, sg, rg), (Io, Vo, so, ro) in zip(a, b):
        assert np.abs(rg - ro).max() <= 1e-13 * np.abs(ro).max()
        assert np.abs(Ig - Io).max() <= 1e-11 * np.abs(Io).max()
        assert np.abs(Vg - Vo).max() <= 1e-11 * np.abs(Vo).max()
        if stage2:
            assert abs(sg - so) <= 1e-11 * max(abs(so), 1e-30) + 1e-11 * np.abs(Io).max()


def test_larger_mesh_transient_against_oracle(require_gpu):
    """N = 938 (1 x 0.6 x 0.6 m box, 10 x 6 x 6 cells), trapezoidal, 20 steps
    against the oracle within 1e-10, and the node potentials and charges of
    transient_recover_node_state against P diag(1/P_aa) V and V / P_jj."""
    L, R, A, P = mesh_operators(*meshgen.box(1.0, 0.6, 0.6, 10, 6, 6))
    ne, nv = A.shape
    gpu = gpu_make(L, R, A, P, TR, 5e-10)
    ora = ck.OracleTransient(L, R, A, P, TR, 5e-10)
    try:
        a = drive(gpu, 20, ne, nv, 11, False)
        b = drive(ora, 20, ne, nv, 11, False)
        V = a[-1][1]
        phi, q = np.zeros(nv), np.zeros(nv)
        lib = plib.load()
        assert lib.transient_recover_node_state(C.byref(gpu.sys), V.ctypes.data_as(ck._dp),
                                                phi.ctypes.data_as(ck._dp), q.ctypes.data_as(ck._dp)) == 0
    finally:
        gpu.close()
    for (Ig, Vg, _, _), (Io, Vo, _, _) in zip(a, b):
        assert np.abs(Ig - Io).max() <= 1e-10 * np.abs(Io).max()
        assert np.abs(Vg - Vo).max() <= 1e-10 * np.abs(Vo).max()
    assert np.array_equal(q, V / np.diag(P))
    want = P @ (V / np.diag(P))
    assert np.abs(phi - want).max() <= 1e-13 * np.abs(want).max()


def test_incidence_operators(require_gpu):
    """incidence_build_AS (equal to the oracle's A S bit for bit),
    incidence_build_APAT (K = A P A^T, per entry as incidence.c forms it),
    incidence_check_row_sum, potential_build_self_capacitance."""
    xyz, quads = meshgen.box(1.0, 0.5, 0.5, 4, 2, 2)
    L, R, A, P = mesh_operators(xyz, quads)
    ne, nv = A.shape
    lib = plib.load()
    Ah = np.ascontiguousarray(A)
    Ph = np.ascontiguousarray(P)
    Am = api.IncidenceMatrix(rows=ne, cols=nv, data=Ah.ctypes.data_as(ck._dp))
    Pm = api.PotentialMatrix(n=nv, data=Ph.ctypes.data_as(ck._dp))
    AS = api.Matrix()
    assert lib.incidence_build_AS(C.byref(Am), C.byref(Pm), C.byref(AS), 1) == 0
    got = api.as_array(AS.data, ne * nv).reshape(ne, nv).copy()
    lib.matrix_free(C.byref(AS))
    _, F, AS_o, _, _ = ck.oracle_transient(L, R, A, P, TR, 1e-9)
    assert np.array_equal(got, AS_o)
    assert lib.incidence_check_row_sum(C.byref(Am)) == 0.0
    Fg = np.zeros(nv)
    assert lib.potential_build_self_capacitance(C.byref(Pm), Fg.ctypes.data_as(ck._dp)) == 0
    assert np.array_equal(Fg, F)
    with plib.Model(xyz, quads) as m:
        K = api.Matrix()
        assert lib.incidence_build_APAT(C.byref(m.mesh), C.byref(Pm), C.byref(K), 1) == 0
        Kg = api.as_array(K.data, ne * ne).reshape(ne, ne).copy()
        lib.matrix_free(C.byref(K))
        e = api.as_array(m.mesh.edges, 2 * ne, np.int32).reshape(-1, 2)
    a, b = e[:, 0], e[:, 1]
    full = P[a][:, a] - P[a][:, b] - P[b][:, a] + P[b][:, b]
    want = np.triu(full) + np.triu(full, 1).T  # upper triangle, mirrored
    assert np.array_equal(Kg, want)
    # a corrupted incidence row is rejected
    bad = Ah.copy()
    bad[3, :] = 0.0
    Bm = api.IncidenceMatrix(rows=ne, cols=nv, data=bad.ctypes.data_as(ck._dp))
    AS2 = api.Matrix()
    assert lib.incidence_build_AS(C.byref(Bm), C.byref(Pm), C.byref(AS2), 1) == -1
