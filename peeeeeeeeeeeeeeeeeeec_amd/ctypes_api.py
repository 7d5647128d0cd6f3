"""ctypes mirrors of the drop-in C structs and the product-library loader.

The struct layouts follow the reference headers byte for byte
(mesh.h:13-133, dual_mesh.h:25-89, complex_matrix.h:21-40, inductance.h,
potential.h, lu.h, complex_lu.h, harmonic.h); the same classes are used to
talk to the GPU library, so a Mesh built here can be handed to any library
exporting that C ABI.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_NAME = "libpeec_mi355x.so"
LIB_PATH = os.path.join(_HERE, LIB_NAME)

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Mesh(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_size_t), ("n_quads", C.c_size_t), ("n_edges", C.c_size_t),
        ("xyz", _dp), ("quads", _ip), ("edges", _ip), ("cell_edges", _ip),
        ("cell_edge_signs", _ip), ("edge_vectors", _dp), ("edge_lengths", _dp),
        ("edge_directions", _dp), ("edge_centers", _dp), ("cell_centers", _dp),
        ("cell_areas", _dp), ("cell_normals", _dp),
    ]


class QuadPatch(C.Structure):
    _fields_ = [("points", C.c_double * 12), ("owner_id", C.c_int), ("cell_id", C.c_int)]


class PatchList(C.Structure):
    _fields_ = [("patches", C.POINTER(QuadPatch)), ("count", C.c_size_t),
                ("capacity", C.c_size_t)]


class DualMesh(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_size_t), ("n_edges", C.c_size_t),
        ("node_patches", C.POINTER(PatchList)), ("edge_patches", C.POINTER(PatchList)),
        ("node_region_areas", _dp), ("edge_region_areas", _dp),
        ("node_region_centers", _dp), ("edge_region_centers", _dp),
    ]


class Complex(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


_cp = C.POINTER(Complex)


class ComplexMatrix(C.Structure):
    _fields_ = [("rows", C.c_size_t), ("cols", C.c_size_t), ("data", _cp)]


class Matrix(C.Structure):
    _fields_ = [("rows", C.c_size_t), ("cols", C.c_size_t), ("data", _dp)]


class InductanceOptions(C.Structure):
    _fields_ = [("far_order", C.c_int), ("near_order", C.c_int),
                ("touching_order", C.c_int), ("self_order", C.c_int),
                ("near_factor", C.c_double), ("parallel_threads", C.c_int),
                ("direction_tolerance", C.c_double)]


class InductanceMatrix(C.Structure):
    _fields_ = [("n", C.c_size_t), ("data", _dp)]


class PotentialOptions(C.Structure):
    _fields_ = [("far_order", C.c_int), ("near_order", C.c_int),
                ("touching_order", C.c_int), ("self_order", C.c_int),
                ("near_factor", C.c_double), ("parallel_threads", C.c_int)]


class PotentialMatrix(C.Structure):
    _fields_ = [("n", C.c_size_t), ("data", _dp)]


class IncidenceMatrix(C.Structure):
    _fields_ = [("rows", C.c_size_t), ("cols", C.c_size_t), ("data", _dp)]


class ResistanceOptions(C.Structure):
    _fields_ = [("model", C.c_int), ("sheet_resistance", C.c_double),
                ("parallel_threads", C.c_int)]


class ResistanceMatrix(C.Structure):
    _fields_ = [("n", C.c_size_t), ("diagonal", _dp)]


class LUOptions(C.Structure):
    _fields_ = [("parallel_threads", C.c_int), ("pivot_tolerance", C.c_double),
                ("parallel_threshold", C.c_size_t)]


class LUFactorization(C.Structure):
    _fields_ = [("n", C.c_size_t), ("data", _dp), ("pivots", C.POINTER(C.c_size_t)),
                ("swap_count", C.c_size_t), ("matrix_scale", C.c_double),
                ("factorized", C.c_int)]


class ComplexLUOptions(LUOptions):
    pass


class ComplexLUFactorization(C.Structure):
    _fields_ = [("n", C.c_size_t), ("data", _cp), ("pivots", C.POINTER(C.c_size_t)),
                ("swap_count", C.c_size_t), ("matrix_scale", C.c_double),
                ("row_scale", _dp), ("column_scale", _dp), ("factorized", C.c_int)]


class InternalCircuitShunt(C.Structure):
    """internal_circuit.h"""
    _fields_ = [("type", C.c_int), ("node_a", C.c_size_t), ("node_b", C.c_size_t), ("resistance", C.c_double),
                ("requested_a", C.c_double * 3), ("requested_b", C.c_double * 3), ("mapped_a", C.c_double * 3),
                ("mapped_b", C.c_double * 3), ("mapping_error_a", C.c_double), ("mapping_error_b", C.c_double),
                ("nodes_mapped", C.c_int), ("initialized", C.c_int)]


class TransientSystem(C.Structure):
    """transient.h"""
    _fields_ = [("stage", C.c_int), ("scheme", C.c_int), ("n_edges", C.c_size_t), ("n_nodes", C.c_size_t),
                ("n_total", C.c_size_t), ("dt", C.c_double), ("L", C.c_void_p), ("R", C.c_void_p),
                ("A", C.c_void_p), ("P", C.c_void_p), ("F", _dp), ("AS", Matrix), ("shunt", C.c_void_p),
                ("b_R", _dp), ("shunt_voltage_operator", _dp), ("slot_circuit", C.c_void_p),
                ("n_slot_cells", C.c_size_t), ("slot_mass_matrix", _dp), ("slot_voltage_operators", _dp),
                ("system_matrix", _dp), ("rhs", _dp), ("solution", _dp), ("lu", LUFactorization),
                ("factorized", C.c_int), ("initialized", C.c_int)]


class HarmonicSystem(C.Structure):
    _fields_ = [("n_edges", C.c_size_t), ("n_nodes", C.c_size_t),
                ("system_size", C.c_size_t), ("frequency", C.c_double),
                ("omega", C.c_double), ("matrix", ComplexMatrix),
                ("lu", ComplexLUFactorization), ("matrix_built", C.c_int),
                ("factorized", C.c_int)]


class HarmonicExcitation(C.Structure):
    _fields_ = [("n_edges", C.c_size_t), ("n_nodes", C.c_size_t),
                ("edge_voltage", _cp), ("node_current", _cp)]


class HarmonicSolution(C.Structure):
    _fields_ = [("n_edges", C.c_size_t), ("n_nodes", C.c_size_t),
                ("edge_current", _cp), ("node_voltage", _cp)]


class HarmonicResidual(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "absolute_residual", "rhs_relative_residual", "backward_error",
        "edge_absolute_residual", "node_absolute_residual",
        "edge_backward_error", "node_backward_error")]


# ---------------------------------------------------------------------------
# helpers shared by every library that speaks this ABI

_libc = C.CDLL(None)
_libc.malloc.restype = C.c_void_p
_libc.malloc.argtypes = [C.c_size_t]


class IncidentField(C.Structure):
    """incident_field.h"""
    _fields_ = [("wave_type", C.c_int), ("polarization", C.c_int), ("theta_deg", C.c_double),
                ("phi_deg", C.c_double), ("theta", C.c_double), ("phi", C.c_double),
                ("k_hat", C.c_double * 3), ("e_theta", C.c_double * 3), ("e_phi", C.c_double * 3),
                ("E_amp", C.c_double), ("E0", C.c_double * 3), ("frequency", C.c_double),
                ("omega", C.c_double), ("wave_number", C.c_double), ("phase", C.c_double),
                ("pulse_k", C.c_double), ("pulse_alpha", C.c_double), ("pulse_beta", C.c_double)]


class ScatteringOptions(C.Structure):
    _fields_ = [("integration", C.c_int), ("parallel_threads", C.c_int)]


class RcsOptions(C.Structure):
    _fields_ = [("integration", C.c_int), ("surface_order", C.c_int), ("parallel_threads", C.c_int)]


class RcsResult(C.Structure):
    _fields_ = [("theta_deg", C.c_double), ("phi_deg", C.c_double), ("direction", C.c_double * 3),
                ("radiation_vector", Complex * 3), ("transverse_vector", Complex * 3), ("sigma", C.c_double),
                ("sigma_dbsm", C.c_double)]


class LightningSource(C.Structure):
    """lightning.h"""
    _fields_ = [("wave_type", C.c_int), ("K", C.c_double), ("alpha", C.c_double), ("beta", C.c_double),
                ("delay", C.c_double), ("strike_node", C.c_size_t), ("return_node", C.c_size_t)]


def c_malloc_copy(arr: np.ndarray, ctype):
    """malloc() a copy of a numpy array (the C side frees it with free())."""
    arr = np.ascontiguousarray(arr)
    ptr = _libc.malloc(max(arr.nbytes, 1))
    if not ptr:
        raise MemoryError("malloc failed")
    C.memmove(ptr, arr.ctypes.data, arr.nbytes)
    return C.cast(ptr, C.POINTER(ctype))


def make_mesh(xyz: np.ndarray, quads: np.ndarray) -> Mesh:
    """A Mesh with malloc'd xyz/quads, topology not yet built."""
    xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    quads = np.asarray(quads, dtype=np.int32).reshape(-1, 4)
    m = Mesh()
    m.n_nodes = xyz.shape[0]
    m.n_quads = quads.shape[0]
    m.n_edges = 0
    m.xyz = c_malloc_copy(xyz, C.c_double)
    m.quads = c_malloc_copy(quads, C.c_int)
    return m


def as_array(ptr, count, dtype=np.float64):
    """Copy `count` elements behind a ctypes pointer into a numpy array."""
    if not ptr or count == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(count,)).astype(dtype, copy=True)


def complex_array(ptr, count):
    if not ptr or count == 0:
        return np.zeros(0, dtype=np.complex128)
    raw = np.ctypeslib.as_array(C.cast(ptr, _dp), shape=(2 * count,))
    return (raw[0::2] + 1j * raw[1::2]).astype(np.complex128)


def complex_buffer(values: np.ndarray):
    """A ctypes Complex array holding `values` (kept alive by the caller)."""
    values = np.ascontiguousarray(np.asarray(values, dtype=np.complex128))
    buf = (Complex * max(values.size, 1))()
    C.memmove(buf, values.ctypes.data, values.nbytes)
    return buf


def mesh_arrays(m: Mesh) -> dict:
    ne, nq, nv = m.n_edges, m.n_quads, m.n_nodes
    return {
        "xyz": as_array(m.xyz, 3 * nv).reshape(-1, 3),
        "quads": as_array(m.quads, 4 * nq, np.int32).reshape(-1, 4),
        "edges": as_array(m.edges, 2 * ne, np.int32).reshape(-1, 2),
        "cell_edges": as_array(m.cell_edges, 4 * nq, np.int32).reshape(-1, 4),
        "cell_edge_signs": as_array(m.cell_edge_signs, 4 * nq, np.int32).reshape(-1, 4),
        "edge_lengths": as_array(m.edge_lengths, ne),
        "edge_directions": as_array(m.edge_directions, 3 * ne).reshape(-1, 3),
        "edge_vectors": as_array(m.edge_vectors, 3 * ne).reshape(-1, 3),
        "edge_centers": as_array(m.edge_centers, 3 * ne).reshape(-1, 3),
        "cell_areas": as_array(m.cell_areas, nq),
    }


def dual_arrays(d: DualMesh) -> dict:
    def plist(lists, n):
        out = []
        for i in range(n):
            pl = lists[i]
            out.append(np.array([[pl.patches[k].points[j] for j in range(12)]
                                 for k in range(pl.count)]).reshape(-1, 12))
        return out
    return {
        "node_patches": plist(d.node_patches, d.n_nodes),
        "edge_patches": plist(d.edge_patches, d.n_edges),
        "node_region_areas": as_array(d.node_region_areas, d.n_nodes),
        "edge_region_areas": as_array(d.edge_region_areas, d.n_edges),
    }
