"""Python host mirror of the drop-in C ABI (libpeec_mi355x.so).

Function names and argument meaning follow the reference's C interface
(inductance.h, potential.h, quadrature.h, incidence.h, lu.h, complex_lu.h,
harmonic.h) so the tests read like the reference's own; results come back as
numpy arrays. There is no fallback: if the HIP library is missing the import
raises, and every call that needs a GPU fails with the library's error.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import ctypes_api as api

_dp = C.POINTER(C.c_double)
_cp = C.POINTER(api.Complex)
_ip = C.POINTER(C.c_int)
_M = C.POINTER(api.Mesh)
_D = C.POINTER(api.DualMesh)
_Q = C.POINTER(api.QuadPatch)


class PeecEngineOptions(C.Structure):
    _fields_ = [("inductance", api.InductanceOptions), ("potential", api.PotentialOptions),
                ("retarded", C.c_int), ("smooth_order", C.c_int), ("near_smooth_order", C.c_int),
                ("block_size", C.c_int),
                ("max_refinement_steps", C.c_int), ("refinement_target", C.c_double),
                ("backward_error_gate", C.c_double), ("reuse_pivots", C.c_int), ("smooth_moments", C.c_int),
                ("shifted_moments", C.c_int), ("split_fp16_update", C.c_int), ("panel_lds_kib", C.c_int),
                ("gmres_krylov", C.c_int)]


class PeecPointReport(C.Structure):
    _fields_ = [("frequency", C.c_double), ("backward_error", C.c_double),
                ("refinement_steps", C.c_int), ("used_fp64_fallback", C.c_int),
                ("pivot_reuse", C.c_int), ("pivot_refactor", C.c_int), ("reuse_backward_error", C.c_double),
                ("ms_assembly", C.c_double), ("ms_fill", C.c_double), ("ms_factor", C.c_double),
                ("ms_solve", C.c_double), ("ms_check", C.c_double), ("factor_cached", C.c_int),
                ("moment_terms", C.c_int), ("block_size", C.c_int), ("moment_center", C.c_double),
                ("batch_size", C.c_int), ("update_tile_rows", C.c_int), ("gmres_steps", C.c_int),
                ("gmres_iterations", C.c_int)]


class PeecServerOptions(C.Structure):
    _fields_ = [("engine", PeecEngineOptions), ("slots", C.c_int), ("hbm_budget_bytes", C.c_size_t),
                ("max_batch", C.c_int)]


class PeecRequestResult(C.Structure):
    _fields_ = [("status", C.c_int), ("report", PeecPointReport), ("queue_s", C.c_double),
                ("latency_s", C.c_double), ("engine", C.c_int)]


class PeecRcsOptions(C.Structure):
    _fields_ = [("wave_type", C.c_int), ("polarization", C.c_int), ("E_amp", C.c_double), ("phase", C.c_double),
                ("excitation", C.c_int), ("rcs", api.RcsOptions)]


class PeecKernelStats(C.Structure):
    _fields_ = [("launches", C.c_longlong), ("total_ms", C.c_double), ("flops", C.c_double),
                ("bytes", C.c_double)]


KERNELS = {"lu_update": 0, "panel": 1, "smooth_tile": 2, "far_tile": 3, "smooth_moments": 4,
           "lu_update_split16": 5, "lu_update_thin": 6, "near_entry": 7, "lu_update_merged": 8}

# flop model of the static assembly: 1/R evaluation, analytic triangle potential
FLOPS_PER_INV_R = 12.0
FLOPS_PER_TRIANGLE_POTENTIAL = 140.0


class PeecStaticStats(C.Structure):
    _fields_ = [("far_evaluations", C.c_longlong), ("near_potential_evaluations", C.c_longlong),
                ("near_pair_evaluations", C.c_longlong), ("far_ms", C.c_double), ("near_ms", C.c_double)]


class RetardationOptions(C.Structure):
    _fields_ = [("inductance", api.InductanceOptions), ("potential", api.PotentialOptions),
                ("smooth_order", C.c_int), ("near_smooth_order", C.c_int)]


# every symbol the public headers under include/ declare (checked by the CPU
# test suite against the built library)
EXPORTED = [
    # mesh.h / dual_mesh.h
    "mesh_init", "mesh_free", "mesh_load_gmsh", "mesh_build_topology", "mesh_print_info",
    "dual_mesh_init", "dual_mesh_free", "dual_mesh_build", "dual_mesh_print_info",
    # quadrature.h
    "gauss_legendre_rule", "triangle_area", "quad_patch_area", "quad_patch_size",
    "quad_patch_min_vertex_distance", "quad_patch_classify_pair", "quad_patch_kernel_1_over_r",
    # inductance.h / potential.h
    "inductance_options_default", "inductance_init", "inductance_free", "inductance_compute_pair",
    "inductance_compute_matrix", "inductance_print_info",
    "potential_options_default", "potential_init", "potential_free", "potential_compute_pair",
    "potential_compute_matrix", "potential_print_info", "potential_build_self_capacitance",
    "potential_self_voltage_to_charge", "potential_self_voltage_to_full_potential",
    # retardation.h
    "retardation_options_default", "retardation_compute_matrices",
    # incidence.h / resistance.h
    "incidence_init", "incidence_free", "incidence_build", "incidence_build_PAT", "incidence_build_AS",
    "incidence_build_APAT", "incidence_check_row_sum", "incidence_print_info",
    # internal_circuit.h / transient.h
    "internal_circuit_shunt_default", "internal_circuit_shunt_set_nodes", "internal_circuit_shunt_map_coordinates",
    "internal_circuit_find_nearest_node", "internal_circuit_shunt_validate",
    "internal_circuit_build_incidence_vector", "internal_circuit_shunt_voltage",
    "internal_circuit_shunt_current_from_voltage", "internal_circuit_shunt_residual",
    "internal_circuit_shunt_print_info",
    "transient_system_init", "transient_system_free", "transient_system_create_stage1",
    "transient_system_create_stage2", "transient_build_shunt_operators", "transient_build_system_matrix",
    "transient_slot_current_offset", "transient_system_create_slot_cells", "transient_build_rhs_slot_cells",
    "transient_extract_slot_cells_solution", "transient_factorize", "transient_build_rhs_stage1",
    "transient_build_rhs_stage2", "transient_solve_step", "transient_current_offset", "transient_voltage_offset",
    "transient_shunt_index", "transient_extract_stage1_solution", "transient_extract_stage2_solution",
    "transient_recover_node_state", "transient_print_info",
    "resistance_options_default", "resistance_init", "resistance_free", "resistance_compute_matrix",
    "resistance_get",
    # matrix.h / complex_matrix.h
    "matrix_init", "matrix_alloc", "matrix_free", "complex_matrix_init", "complex_matrix_alloc",
    "complex_matrix_free", "complex_matrix_vector_multiply",
    # lu.h / complex_lu.h
    "lu_options_default", "lu_init", "lu_free", "lu_factorize", "lu_solve", "lu_solve_multiple",
    "lu_determinant", "lu_print_info",
    "complex_lu_options_default", "complex_lu_init", "complex_lu_free", "complex_lu_factorize",
    "complex_lu_solve", "complex_lu_solve_multiple", "complex_lu_print_info",
    # harmonic.h
    "harmonic_system_init", "harmonic_system_free", "harmonic_solution_init",
    "harmonic_solution_free", "harmonic_build_system", "harmonic_factorize", "harmonic_build_rhs",
    "harmonic_solve", "harmonic_compute_residual", "harmonic_compute_residual_noalloc", "harmonic_print_info",
    # peec_gpu.h
    "peec_engine_options_default", "peec_engine_create", "peec_engine_create_with_matrices", "peec_engine_solve",
    "peec_engine_solve_batch",
    "peec_engine_static_matrices", "peec_engine_retarded_matrices", "peec_engine_edges",
    "peec_engine_nodes", "peec_engine_far_evaluations", "peec_engine_smooth_evaluations",
    "peec_engine_destroy", "peec_engine_set_profiling", "peec_engine_kernel_stats", "peec_engine_static_stats",
    "peec_device_count", "peec_device_synchronize", "peec_selftest_mfma_layout", "peec_rcs_options_default", "peec_engine_monostatic_rcs",
    # incident_field.h / scattering.h / rcs.h
    "incident_field_init", "incident_field_configure_harmonic", "incident_field_configure_pulse",
    "incident_field_evaluate_time", "incident_field_evaluate_phasor", "incident_field_print_info",
    "scattering_options_default", "scattering_build_harmonic_excitation", "scattering_build_time_excitation",
    "scattering_zero_node_current_real", "scattering_zero_node_current_complex",
    "scattering_print_harmonic_excitation_info", "scattering_print_time_excitation_info",
    "rcs_options_default", "rcs_build_observation_direction", "rcs_compute_radiation_vector",
    "rcs_compute_direction", "rcs_compute_far_electric_field", "rcs_print_result",
    # lightning.h and the frequency-domain transient (peec_gpu.h)
    "lightning_source_default", "lightning_source_validate", "lightning_current", "lightning_peak",
    "lightning_build_node_current", "lightning_find_nearest_node", "lightning_print_info",
    "peec_engine_port_sweep", "peec_lightning_spectrum", "peec_port_transient_from_spectrum",
    "peec_engine_lightning_response",
    # matrix_export.h
    "matrix_export_dense_mtx", "matrix_export_incidence_mtx", "matrix_export_resistance_mtx",
    "matrix_export_peec_set", "matrix_import_dense_mtx",
    # peec_serve.h
    "peec_server_options_default", "peec_server_create", "peec_server_add_mesh", "peec_server_submit",
    "peec_server_wait", "peec_server_memory", "peec_server_destroy", "peec_server_engine_bytes",
]


def _sig(lib, name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)


_lib = None


def load():
    """Load the in-tree HIP library; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(api.LIB_PATH):
        raise ImportError(
            f"{api.LIB_PATH} is missing: build it with `python -m peeeeeeeeeeeeeeeeeeec_amd.build` "
            "(there is no CPU fallback)")
    lib = C.CDLL(api.LIB_PATH)
    _sig(lib, "mesh_init", None, _M)
    _sig(lib, "mesh_free", None, _M)
    _sig(lib, "mesh_load_gmsh", C.c_int, C.c_char_p, _M)
    _sig(lib, "mesh_build_topology", C.c_int, _M)
    _sig(lib, "dual_mesh_init", None, _D)
    _sig(lib, "dual_mesh_free", None, _D)
    _sig(lib, "dual_mesh_build", C.c_int, _M, _D)
    _sig(lib, "gauss_legendre_rule", C.c_int, C.c_int, _dp, _dp)
    _sig(lib, "quad_patch_classify_pair", C.c_int, _Q, _Q, C.c_double)
    _sig(lib, "quad_patch_kernel_1_over_r", C.c_double, _Q, _Q, C.c_int, C.c_int, C.c_int,
         C.c_int, C.c_double)
    _sig(lib, "quad_patch_area", C.c_double, _Q)
    _sig(lib, "quad_patch_size", C.c_double, _Q)
    _sig(lib, "inductance_options_default", None, C.POINTER(api.InductanceOptions), C.c_int)
    _sig(lib, "potential_options_default", None, C.POINTER(api.PotentialOptions), C.c_int)
    _sig(lib, "inductance_compute_matrix", C.c_int, _M, _D, C.POINTER(api.InductanceOptions),
         C.POINTER(api.InductanceMatrix))
    _sig(lib, "potential_compute_matrix", C.c_int, _M, _D, C.POINTER(api.PotentialOptions),
         C.POINTER(api.PotentialMatrix))
    _sig(lib, "inductance_compute_pair", C.c_int, _M, _D, C.c_size_t, C.c_size_t,
         C.POINTER(api.InductanceOptions), _dp)
    _sig(lib, "potential_compute_pair", C.c_int, _M, _D, C.c_size_t, C.c_size_t,
         C.POINTER(api.PotentialOptions), _dp)
    _sig(lib, "inductance_free", None, C.POINTER(api.InductanceMatrix))
    _sig(lib, "potential_free", None, C.POINTER(api.PotentialMatrix))
    _sig(lib, "retardation_options_default", None, C.POINTER(RetardationOptions), C.c_int)
    _sig(lib, "retardation_compute_matrices", C.c_int, _M, _D, C.c_double,
         C.POINTER(RetardationOptions), C.POINTER(api.ComplexMatrix), C.POINTER(api.ComplexMatrix))
    _sig(lib, "complex_matrix_free", None, C.POINTER(api.ComplexMatrix))
    _sig(lib, "matrix_free", None, C.POINTER(api.Matrix))
    _sig(lib, "incidence_build", C.c_int, _M, C.POINTER(api.IncidenceMatrix))
    _sig(lib, "incidence_free", None, C.POINTER(api.IncidenceMatrix))
    _sig(lib, "incidence_build_PAT", C.c_int, _M, C.POINTER(api.PotentialMatrix),
         C.POINTER(api.Matrix), C.c_int)
    _sig(lib, "resistance_options_default", None, C.POINTER(api.ResistanceOptions), C.c_int)
    _sig(lib, "resistance_compute_matrix", C.c_int, _M, _D, C.POINTER(api.ResistanceOptions),
         C.POINTER(api.ResistanceMatrix))
    _sig(lib, "resistance_free", None, C.POINTER(api.ResistanceMatrix))
    _sig(lib, "complex_matrix_vector_multiply", C.c_int, C.POINTER(api.ComplexMatrix), _cp, _cp, C.c_int)
    _sig(lib, "lu_options_default", None, C.POINTER(api.LUOptions), C.c_int)
    _sig(lib, "lu_init", None, C.POINTER(api.LUFactorization))
    _sig(lib, "lu_free", None, C.POINTER(api.LUFactorization))
    _sig(lib, "lu_factorize", C.c_int, _dp, C.c_size_t, C.POINTER(api.LUOptions),
         C.POINTER(api.LUFactorization))
    _sig(lib, "lu_solve", C.c_int, C.POINTER(api.LUFactorization), _dp, _dp)
    _PM, _IM, _MX = C.POINTER(api.PotentialMatrix), C.POINTER(api.IncidenceMatrix), C.POINTER(api.Matrix)
    _SH, _TS = C.POINTER(api.InternalCircuitShunt), C.POINTER(api.TransientSystem)
    _sig(lib, "potential_build_self_capacitance", C.c_int, _PM, _dp)
    _sig(lib, "potential_self_voltage_to_charge", C.c_int, _PM, _dp, _dp)
    _sig(lib, "potential_self_voltage_to_full_potential", C.c_int, _PM, _dp, _dp)
    _sig(lib, "incidence_build_AS", C.c_int, _IM, _PM, _MX, C.c_int)
    _sig(lib, "incidence_build_APAT", C.c_int, _M, _PM, _MX, C.c_int)
    _sig(lib, "incidence_check_row_sum", C.c_double, _IM)
    _sig(lib, "internal_circuit_shunt_default", None, _SH)
    _sig(lib, "internal_circuit_shunt_set_nodes", C.c_int, _SH, _M, C.c_size_t, C.c_size_t)
    _sig(lib, "internal_circuit_shunt_map_coordinates", C.c_int, _SH, _M, _dp, _dp)
    _sig(lib, "internal_circuit_find_nearest_node", C.c_int, _M, _dp, C.POINTER(C.c_size_t), _dp)
    _sig(lib, "internal_circuit_shunt_validate", C.c_int, _SH, C.c_size_t)
    _sig(lib, "internal_circuit_build_incidence_vector", C.c_int, _SH, C.c_size_t, _dp)
    _sig(lib, "internal_circuit_shunt_voltage", C.c_double, _SH, _dp, C.c_size_t)
    _sig(lib, "internal_circuit_shunt_current_from_voltage", C.c_double, _SH, _dp, C.c_size_t)
    _sig(lib, "internal_circuit_shunt_residual", C.c_double, _SH, _dp, C.c_size_t, C.c_double)
    _sig(lib, "transient_system_init", None, _TS)
    _sig(lib, "transient_system_free", None, _TS)
    _sig(lib, "transient_system_create_stage1", C.c_int, _TS, C.POINTER(api.InductanceMatrix),
         C.POINTER(api.ResistanceMatrix), _IM, _PM, C.c_int, C.c_double)
    _sig(lib, "transient_system_create_stage2", C.c_int, _TS, C.POINTER(api.InductanceMatrix),
         C.POINTER(api.ResistanceMatrix), _IM, _PM, _SH, C.c_int, C.c_double)
    _sig(lib, "transient_build_system_matrix", C.c_int, _TS)
    _sig(lib, "transient_factorize", C.c_int, _TS, C.c_int)
    _sig(lib, "transient_build_rhs_stage1", C.c_int, _TS, _dp, _dp, _dp, _dp, _dp, _dp)
    _sig(lib, "transient_build_rhs_stage2", C.c_int, _TS, _dp, _dp, C.c_double, _dp, _dp, _dp, _dp)
    _sig(lib, "transient_solve_step", C.c_int, _TS)
    _sig(lib, "transient_shunt_index", C.c_size_t, _TS)
    _sig(lib, "transient_extract_stage1_solution", C.c_int, _TS, _dp, _dp)
    _sig(lib, "transient_extract_stage2_solution", C.c_int, _TS, _dp, _dp, _dp)
    _sig(lib, "transient_recover_node_state", C.c_int, _TS, _dp, _dp, _dp)
    _sig(lib, "lu_solve_multiple", C.c_int, C.POINTER(api.LUFactorization), _dp, _dp, C.c_size_t, C.c_int)
    _sig(lib, "lu_determinant", C.c_double, C.POINTER(api.LUFactorization))
    _sig(lib, "complex_lu_options_default", None, C.POINTER(api.ComplexLUOptions), C.c_int)
    _sig(lib, "complex_lu_init", None, C.POINTER(api.ComplexLUFactorization))
    _sig(lib, "complex_lu_free", None, C.POINTER(api.ComplexLUFactorization))
    _sig(lib, "complex_lu_factorize", C.c_int, _cp, C.c_size_t, C.POINTER(api.ComplexLUOptions),
         C.POINTER(api.ComplexLUFactorization))
    _sig(lib, "complex_lu_solve", C.c_int, C.POINTER(api.ComplexLUFactorization), _cp, _cp)
    _sig(lib, "complex_lu_solve_multiple", C.c_int, C.POINTER(api.ComplexLUFactorization), _cp, _cp,
         C.c_size_t, C.c_int)
    _sig(lib, "harmonic_system_init", None, C.POINTER(api.HarmonicSystem))
    _sig(lib, "harmonic_system_free", None, C.POINTER(api.HarmonicSystem))
    _sig(lib, "harmonic_solution_init", None, C.POINTER(api.HarmonicSolution))
    _sig(lib, "harmonic_solution_free", None, C.POINTER(api.HarmonicSolution))
    _sig(lib, "harmonic_build_system", C.c_int, C.POINTER(api.InductanceMatrix),
         C.POINTER(api.ResistanceMatrix), C.POINTER(api.IncidenceMatrix), C.POINTER(api.Matrix),
         C.c_double, C.c_int, C.POINTER(api.HarmonicSystem))
    _sig(lib, "harmonic_factorize", C.c_int, C.POINTER(api.HarmonicSystem), C.c_int)
    _sig(lib, "harmonic_build_rhs", C.c_int, C.POINTER(api.PotentialMatrix),
         C.POINTER(api.HarmonicExcitation), _cp, C.c_int)
    _sig(lib, "harmonic_solve", C.c_int, C.POINTER(api.HarmonicSystem), C.POINTER(api.PotentialMatrix),
         C.POINTER(api.HarmonicExcitation), C.POINTER(api.HarmonicSolution), C.c_int)
    _sig(lib, "harmonic_compute_residual", C.c_int, C.POINTER(api.HarmonicSystem),
         C.POINTER(api.PotentialMatrix), C.POINTER(api.HarmonicExcitation),
         C.POINTER(api.HarmonicSolution), C.c_int, C.POINTER(api.HarmonicResidual))
    _sig(lib, "harmonic_compute_residual_noalloc", C.c_int, C.POINTER(api.HarmonicSystem),
         C.POINTER(api.PotentialMatrix), C.POINTER(api.HarmonicExcitation),
         C.POINTER(api.HarmonicSolution), C.c_int, C.POINTER(api.HarmonicResidual))
    _sig(lib, "peec_engine_options_default", None, C.POINTER(PeecEngineOptions))
    _sig(lib, "peec_engine_create", C.c_int, _M, _D, C.POINTER(PeecEngineOptions), C.POINTER(C.c_void_p))
    _sig(lib, "peec_engine_create_with_matrices", C.c_int, _M, _D, C.POINTER(PeecEngineOptions), _dp, _dp,
         C.POINTER(C.c_void_p))
    _sig(lib, "peec_engine_solve", C.c_int, C.c_void_p, C.c_double, C.c_int, _ip, C.c_size_t, _cp, _cp,
         _cp, C.POINTER(PeecPointReport))
    _sig(lib, "peec_engine_solve_batch", C.c_int, C.c_void_p, C.c_size_t, _dp, C.POINTER(_ip),
         C.POINTER(C.c_size_t), C.POINTER(_cp), C.POINTER(PeecPointReport), _ip)
    _sig(lib, "peec_engine_static_matrices", C.c_int, C.c_void_p, _dp, _dp)
    _sig(lib, "peec_engine_retarded_matrices", C.c_int, C.c_void_p, C.c_double, _cp, _cp)
    _sig(lib, "peec_engine_edges", C.c_size_t, C.c_void_p)
    _sig(lib, "peec_engine_nodes", C.c_size_t, C.c_void_p)
    _sig(lib, "peec_engine_far_evaluations", C.c_longlong, C.c_void_p)
    _sig(lib, "peec_engine_smooth_evaluations", C.c_longlong, C.c_void_p)
    _sig(lib, "peec_engine_destroy", None, C.c_void_p)
    _sig(lib, "peec_server_options_default", None, C.POINTER(PeecServerOptions))
    _sig(lib, "peec_server_create", C.c_int, C.POINTER(PeecServerOptions), C.POINTER(C.c_void_p))
    _sig(lib, "peec_server_add_mesh", C.c_int, C.c_void_p, _M, _D, C.POINTER(C.c_int))
    _sig(lib, "peec_server_submit", C.c_int, C.c_void_p, C.c_int, C.c_double, C.c_int, _ip, C.c_size_t,
         C.POINTER(C.c_longlong))
    _sig(lib, "peec_server_wait", C.c_int, C.c_void_p, C.c_longlong, _cp, C.POINTER(PeecRequestResult))
    _sig(lib, "peec_server_memory", C.c_size_t, C.c_void_p, C.POINTER(C.c_int))
    _sig(lib, "peec_server_destroy", None, C.c_void_p)
    _sig(lib, "peec_server_engine_bytes", C.c_size_t, C.c_size_t, C.c_size_t, C.c_int)
    _sig(lib, "peec_engine_set_profiling", C.c_int, C.c_void_p, C.c_int)
    _sig(lib, "peec_selftest_mfma_layout", C.c_int, _dp)
    _sig(lib, "peec_device_count", C.c_int)
    _sig(lib, "peec_device_synchronize", C.c_int)
    _sig(lib, "peec_engine_kernel_stats", C.c_int, C.c_void_p, C.c_int, C.POINTER(PeecKernelStats))
    _sig(lib, "peec_engine_static_stats", C.c_int, C.c_void_p, C.POINTER(PeecStaticStats))
    _F = C.POINTER(api.IncidentField)
    _sig(lib, "incident_field_configure_harmonic", C.c_int, _F, C.c_int, C.c_int, C.c_double, C.c_double,
         C.c_double, C.c_double, C.c_double)
    _sig(lib, "incident_field_configure_pulse", C.c_int, _F, C.c_int, C.c_double, C.c_double, C.c_double,
         C.c_double, C.c_double)
    _sig(lib, "incident_field_evaluate_phasor", C.c_int, _F, _dp, _cp)
    _sig(lib, "incident_field_evaluate_time", C.c_int, _F, _dp, C.c_double, _dp)
    _sig(lib, "scattering_options_default", None, C.POINTER(api.ScatteringOptions), C.c_int)
    _sig(lib, "scattering_build_harmonic_excitation", C.c_int, _M, _F, C.POINTER(api.ScatteringOptions), _cp)
    _sig(lib, "scattering_build_time_excitation", C.c_int, _M, _F, C.POINTER(api.ScatteringOptions),
         C.c_double, _dp)
    _sig(lib, "rcs_options_default", None, C.POINTER(api.RcsOptions), C.c_int)
    _sig(lib, "rcs_build_observation_direction", C.c_int, C.c_double, C.c_double, _dp)
    _sig(lib, "rcs_compute_radiation_vector", C.c_int, _M, _D, _cp, C.c_double, _dp,
         C.POINTER(api.RcsOptions), _cp)
    _sig(lib, "rcs_compute_direction", C.c_int, _M, _D, _F, C.POINTER(api.HarmonicSolution),
         C.POINTER(api.RcsOptions), C.c_double, C.c_double, C.POINTER(api.RcsResult))
    _sig(lib, "rcs_compute_far_electric_field", C.c_int, C.POINTER(api.RcsResult), C.c_double, C.c_double, _cp)
    _sig(lib, "peec_rcs_options_default", None, C.POINTER(PeecRcsOptions))
    _sig(lib, "peec_engine_monostatic_rcs", C.c_int, C.c_void_p, C.c_double, C.c_int, C.POINTER(PeecRcsOptions),
         _dp, _dp, C.c_size_t, C.POINTER(api.RcsResult), _cp, C.POINTER(PeecPointReport))
    _sig(lib, "matrix_export_dense_mtx", C.c_int, C.c_char_p, _dp, C.c_size_t, C.c_size_t)
    _sig(lib, "matrix_export_incidence_mtx", C.c_int, C.c_char_p, _M)
    _sig(lib, "matrix_export_resistance_mtx", C.c_int, C.c_char_p, C.POINTER(api.ResistanceMatrix))
    _sig(lib, "matrix_export_peec_set", C.c_int, C.c_char_p, _M, C.POINTER(api.InductanceMatrix),
         C.POINTER(api.PotentialMatrix), C.POINTER(api.Matrix), C.POINTER(api.ResistanceMatrix))
    _sig(lib, "matrix_import_dense_mtx", C.c_int, C.c_char_p, C.POINTER(_dp), C.POINTER(C.c_size_t),
         C.POINTER(C.c_size_t))
    _S = C.POINTER(api.LightningSource)
    _sig(lib, "lightning_source_default", None, _S)
    _sig(lib, "lightning_source_validate", C.c_int, _S, C.c_size_t)
    _sig(lib, "lightning_current", C.c_double, _S, C.c_double)
    _sig(lib, "lightning_peak", C.c_int, _S, _dp, _dp)
    _sig(lib, "lightning_build_node_current", C.c_int, _S, C.c_size_t, C.c_double, _dp)
    _sig(lib, "lightning_find_nearest_node", C.c_int, _M, _dp, C.POINTER(C.c_size_t), _dp)
    _sig(lib, "peec_engine_port_sweep", C.c_int, C.c_void_p, _dp, C.c_size_t, C.c_int, _ip, _cp,
         C.POINTER(PeecPointReport))
    _sig(lib, "peec_lightning_spectrum", C.c_int, _S, C.c_double, _cp)
    _sig(lib, "peec_port_transient_from_spectrum", C.c_int, _S, C.c_size_t, C.c_double, _cp, _dp, _dp, _dp)
    _sig(lib, "peec_engine_lightning_response", C.c_int, C.c_void_p, _S, C.c_size_t, C.c_double, C.c_int, _dp,
         _dp, _dp, _cp, C.POINTER(PeecPointReport))
    _lib = lib
    return lib


# ---------------------------------------------------------------------------
# thin numpy-level wrappers


class Model:
    """A Mesh + DualMesh built by the product library from xyz/quads or .msh."""

    def __init__(self, xyz=None, quads=None, msh=None):
        lib = load()
        if msh is not None:
            self.mesh = api.Mesh()
            if lib.mesh_load_gmsh(msh.encode(), C.byref(self.mesh)) != 0:
                raise RuntimeError(f"cannot load {msh}")
        else:
            self.mesh = api.make_mesh(xyz, quads)
        if lib.mesh_build_topology(C.byref(self.mesh)) != 0:
            raise RuntimeError("mesh_build_topology failed")
        self.dual = api.DualMesh()
        if lib.dual_mesh_build(C.byref(self.mesh), C.byref(self.dual)) != 0:
            raise RuntimeError("dual_mesh_build failed")
        self.ne = self.mesh.n_edges
        self.nv = self.mesh.n_nodes

    def close(self):
        lib = load()
        lib.dual_mesh_free(C.byref(self.dual))
        lib.mesh_free(C.byref(self.mesh))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def edges(self):
        return api.as_array(self.mesh.edges, 2 * self.ne, np.int32).reshape(-1, 2)

    def inductance(self, options=None):
        lib = load()
        o = options or api.InductanceOptions()
        if options is None:
            lib.inductance_options_default(C.byref(o), 1)
        L = api.InductanceMatrix()
        st = lib.inductance_compute_matrix(C.byref(self.mesh), C.byref(self.dual), C.byref(o), C.byref(L))
        if st != 0:
            raise RuntimeError(f"inductance_compute_matrix failed ({st})")
        out = api.as_array(L.data, L.n * L.n).reshape(L.n, L.n)
        lib.inductance_free(C.byref(L))
        return out

    def potential(self, options=None):
        lib = load()
        o = options or api.PotentialOptions()
        if options is None:
            lib.potential_options_default(C.byref(o), 1)
        P = api.PotentialMatrix()
        st = lib.potential_compute_matrix(C.byref(self.mesh), C.byref(self.dual), C.byref(o), C.byref(P))
        if st != 0:
            raise RuntimeError(f"potential_compute_matrix failed ({st})")
        out = api.as_array(P.data, P.n * P.n).reshape(P.n, P.n)
        lib.potential_free(C.byref(P))
        return out

    def retarded(self, wave_number, smooth_order=2, near_smooth_order=4):
        lib = load()
        o = RetardationOptions()
        lib.retardation_options_default(C.byref(o), 1)
        o.smooth_order = smooth_order
        o.near_smooth_order = near_smooth_order
        L = api.ComplexMatrix()
        P = api.ComplexMatrix()
        st = lib.retardation_compute_matrices(C.byref(self.mesh), C.byref(self.dual), wave_number,
                                              C.byref(o), C.byref(L), C.byref(P))
        if st != 0:
            raise RuntimeError(f"retardation_compute_matrices failed ({st})")
        Lc = api.complex_array(L.data, L.rows * L.cols).reshape(L.rows, L.cols)
        Pc = api.complex_array(P.data, P.rows * P.cols).reshape(P.rows, P.cols)
        lib.complex_matrix_free(C.byref(L))
        lib.complex_matrix_free(C.byref(P))
        return Lc, Pc


def plane_wave(frequency, theta_deg, phi_deg, polarization=1, E_amp=1.0, phase=0.0, wave_type=2):
    """incident_field_configure_harmonic (wave_type 2 = COMPLEX)."""
    lib = load()
    f = api.IncidentField()
    st = lib.incident_field_configure_harmonic(C.byref(f), wave_type, polarization, E_amp, frequency, theta_deg,
                                               phi_deg, phase)
    if st != 0:
        raise RuntimeError("incident_field_configure_harmonic failed")
    return f


def excitation(model, field, integration=1):
    """scattering_build_harmonic_excitation -> U (Ne complex)."""
    lib = load()
    o = api.ScatteringOptions()
    lib.scattering_options_default(C.byref(o), 1)
    o.integration = integration
    U = np.zeros(model.ne, np.complex128)
    st = lib.scattering_build_harmonic_excitation(C.byref(model.mesh), C.byref(field), C.byref(o),
                                                  U.ctypes.data_as(_cp))
    if st != 0:
        raise RuntimeError("scattering_build_harmonic_excitation failed")
    return U


def radiation_vector(model, currents, wave_number, direction, integration=1, order=6):
    """rcs_compute_radiation_vector -> F (3 complex)."""
    lib = load()
    o = api.RcsOptions()
    lib.rcs_options_default(C.byref(o), 1)
    o.integration, o.surface_order = integration, order
    I = np.ascontiguousarray(currents, np.complex128)
    d = np.ascontiguousarray(direction, np.float64)
    F = np.zeros(3, np.complex128)
    st = lib.rcs_compute_radiation_vector(C.byref(model.mesh), C.byref(model.dual), I.ctypes.data_as(_cp),
                                          wave_number, d.ctypes.data_as(_dp), C.byref(o), F.ctypes.data_as(_cp))
    if st != 0:
        raise RuntimeError("rcs_compute_radiation_vector failed")
    return F


def rcs_direction(model, field, currents, theta_deg, phi_deg, integration=1, order=6):
    """rcs_compute_direction on a solution with the given edge currents."""
    lib = load()
    o = api.RcsOptions()
    lib.rcs_options_default(C.byref(o), 1)
    o.integration, o.surface_order = integration, order
    I = np.ascontiguousarray(currents, np.complex128)
    sol = api.HarmonicSolution()
    sol.n_edges, sol.n_nodes = model.ne, model.nv
    sol.edge_current = C.cast(I.ctypes.data, C.POINTER(api.Complex))
    r = api.RcsResult()
    st = lib.rcs_compute_direction(C.byref(model.mesh), C.byref(model.dual), C.byref(field), C.byref(sol),
                                   C.byref(o), theta_deg, phi_deg, C.byref(r))
    if st != 0:
        raise RuntimeError("rcs_compute_direction failed")
    return r


def lightning_source(K=102900.0, alpha=1500.0, beta=1.0e6, delay=0.0, strike=0, ret=0):
    s = api.LightningSource()
    load().lightning_source_default(C.byref(s))
    s.K, s.alpha, s.beta, s.delay, s.strike_node, s.return_node = K, alpha, beta, delay, strike, ret
    return s


def transient_from_spectrum(source, df, z_in):
    """peec_port_transient_from_spectrum: (time, current, voltage)."""
    z = np.ascontiguousarray(z_in, np.complex128)
    n = z.size
    t, i, v = np.zeros(2 * n), np.zeros(2 * n), np.zeros(2 * n)
    st = load().peec_port_transient_from_spectrum(C.byref(source), n, df, z.ctypes.data_as(_cp),
                                                  t.ctypes.data_as(_dp), i.ctypes.data_as(_dp),
                                                  v.ctypes.data_as(_dp))
    if st != 0:
        raise RuntimeError("peec_port_transient_from_spectrum failed")
    return t, i, v


def complex_lu_factor_solve(Z, B):
    """Factor Z with complex_lu_factorize and solve the columns of B (n x k)."""
    lib = load()
    Z = np.ascontiguousarray(Z, dtype=np.complex128)
    n = Z.shape[0]
    o = api.ComplexLUOptions()
    lib.complex_lu_options_default(C.byref(o), 1)
    lu = api.ComplexLUFactorization()
    st = lib.complex_lu_factorize(Z.ctypes.data_as(_cp), n, C.byref(o), C.byref(lu))
    if st != 0:
        lib.complex_lu_free(C.byref(lu))
        return st, None, None
    Bm = np.ascontiguousarray(np.asarray(B, dtype=np.complex128).reshape(-1, n))
    X = np.zeros_like(Bm)
    st = lib.complex_lu_solve_multiple(C.byref(lu), Bm.ctypes.data_as(_cp), X.ctypes.data_as(_cp),
                                       Bm.shape[0], 1)
    info = {
        "data": api.complex_array(lu.data, n * n).reshape(n, n),
        "pivots": api.as_array(lu.pivots, n, np.uint64).astype(np.int64),
        "row_scale": api.as_array(lu.row_scale, n),
        "column_scale": api.as_array(lu.column_scale, n),
        "matrix_scale": lu.matrix_scale,
        "swap_count": lu.swap_count,
    }
    lib.complex_lu_free(C.byref(lu))
    return st, X, info


def real_lu_factor_solve(A, B, want_factor=False):
    """Factor A with lu_factorize and solve the columns of B; the third result is
    the determinant, or (determinant, factor, pivots) with want_factor."""
    lib = load()
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0]
    o = api.LUOptions()
    lib.lu_options_default(C.byref(o), 1)
    lu = api.LUFactorization()
    st = lib.lu_factorize(A.ctypes.data_as(_dp), n, C.byref(o), C.byref(lu))
    if st != 0:
        lib.lu_free(C.byref(lu))
        return st, None, None
    Bm = np.ascontiguousarray(np.asarray(B, dtype=np.float64).reshape(-1, n))
    X = np.zeros_like(Bm)
    st = lib.lu_solve_multiple(C.byref(lu), Bm.ctypes.data_as(_dp), X.ctypes.data_as(_dp), Bm.shape[0], 1)
    det = lib.lu_determinant(C.byref(lu))
    if want_factor:
        det = (det, api.as_array(lu.data, n * n).reshape(n, n).copy(),
               api.as_array(lu.pivots, n, np.uint64).astype(np.int64))
    lib.lu_free(C.byref(lu))
    return st, X, det


def device_count():
    """Visible HIP devices, through the library's own runtime."""
    return load().peec_device_count()


def device_synchronize():
    """hipDeviceSynchronize on the current device."""
    if load().peec_device_synchronize() != 0:
        raise RuntimeError("peec_device_synchronize failed")


def complex_matvec(A, x):
    """y = A x through complex_matrix_vector_multiply (any rows x cols)."""
    lib = load()
    A = np.ascontiguousarray(A, dtype=np.complex128)
    rows, cols = A.shape
    x = np.ascontiguousarray(x, dtype=np.complex128)
    M = api.ComplexMatrix()
    M.rows, M.cols = rows, cols
    M.data = A.ctypes.data_as(_cp)
    y = np.zeros(rows, np.complex128)
    if lib.complex_matrix_vector_multiply(C.byref(M), x.ctypes.data_as(_cp), y.ctypes.data_as(_cp), 1) != 0:
        raise RuntimeError("complex_matrix_vector_multiply failed")
    return y


class Engine:
    """Device-resident frequency-point engine (include/peec_gpu.h)."""

    FP64 = 0
    MIXED = 1

    def __init__(self, model: Model, retarded=False, smooth_order=2, near_smooth_order=4, block_size=0,
                 static_matrices=None, **opts):
        lib = load()
        self.model = model
        o = PeecEngineOptions()
        lib.peec_engine_options_default(C.byref(o))
        o.retarded = int(bool(retarded))
        o.smooth_order = smooth_order
        o.near_smooth_order = near_smooth_order
        o.block_size = block_size
        for k, v in opts.items():
            setattr(o, k, v)
        self.options = o
        h = C.c_void_p()
        if static_matrices is not None:
            Lh = np.ascontiguousarray(static_matrices[0], np.float64)
            Ph = np.ascontiguousarray(static_matrices[1], np.float64)
            assert Lh.shape == (model.ne, model.ne) and Ph.shape == (model.nv, model.nv)
            st = lib.peec_engine_create_with_matrices(C.byref(model.mesh), C.byref(model.dual), C.byref(o),
                                                      Lh.ctypes.data_as(_dp), Ph.ctypes.data_as(_dp), C.byref(h))
        else:
            st = lib.peec_engine_create(C.byref(model.mesh), C.byref(model.dual), C.byref(o), C.byref(h))
        if st != 0:
            raise RuntimeError(f"peec_engine_create failed ({st})")
        self.handle = h
        self.ne = model.ne
        self.nv = model.nv

    def close(self):
        if self.handle:
            load().peec_engine_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def static_matrices(self):
        L = np.zeros(self.ne * self.ne)
        P = np.zeros(self.nv * self.nv)
        if load().peec_engine_static_matrices(self.handle, L.ctypes.data_as(_dp), P.ctypes.data_as(_dp)) != 0:
            raise RuntimeError("peec_engine_static_matrices failed")
        return L.reshape(self.ne, self.ne), P.reshape(self.nv, self.nv)

    def retarded_matrices(self, k):
        L = np.zeros(self.ne * self.ne, dtype=np.complex128)
        P = np.zeros(self.nv * self.nv, dtype=np.complex128)
        if load().peec_engine_retarded_matrices(self.handle, k, L.ctypes.data_as(_cp), P.ctypes.data_as(_cp)) != 0:
            raise RuntimeError("peec_engine_retarded_matrices failed")
        return L.reshape(self.ne, self.ne), P.reshape(self.nv, self.nv)

    def solve(self, frequency, ports, mode=0, want_solution=True):
        """ports: list of (a, b) node pairs. Returns (status, z_in, I, V, report)."""
        ports = np.ascontiguousarray(np.asarray(ports, dtype=np.int32).reshape(-1, 2))
        k = ports.shape[0]
        z = np.zeros(k, dtype=np.complex128)
        I = np.zeros(k * self.ne, dtype=np.complex128) if want_solution else None
        V = np.zeros(k * self.nv, dtype=np.complex128) if want_solution else None
        rep = PeecPointReport()
        st = load().peec_engine_solve(
            self.handle, float(frequency), int(mode), ports.ctypes.data_as(_ip), k,
            z.ctypes.data_as(_cp), I.ctypes.data_as(_cp) if I is not None else None,
            V.ctypes.data_as(_cp) if V is not None else None, C.byref(rep))
        if I is not None:
            I = I.reshape(k, self.ne)
            V = V.reshape(k, self.nv)
        return st, z, I, V, rep

    def solve_batch(self, frequencies, ports):
        """Mixed-precision points solved together (peec_engine_solve_batch):
        ports[p] is the list of (a, b) node pairs of point p (one list for
        all points if a single list of pairs is given). Returns (status,
        per-point statuses, per-point z_in arrays, reports)."""
        f = np.ascontiguousarray(frequencies, np.float64)
        npt = f.size
        if len(ports) and np.ndim(ports[0]) == 1:
            ports = [ports] * npt
        parr = [np.ascontiguousarray(np.asarray(p, dtype=np.int32).reshape(-1, 2)) for p in ports]
        assert len(parr) == npt
        zs = [np.zeros(p.shape[0], np.complex128) for p in parr]
        pp = (_ip * npt)(*[p.ctypes.data_as(_ip) for p in parr])
        npo = (C.c_size_t * npt)(*[p.shape[0] for p in parr])
        zp = (_cp * npt)(*[z.ctypes.data_as(_cp) for z in zs])
        reps = (PeecPointReport * npt)()
        st = np.zeros(npt, np.int32)
        code = load().peec_engine_solve_batch(self.handle, npt, f.ctypes.data_as(_dp), pp, npo, zp, reps,
                                              st.ctypes.data_as(_ip))
        return code, st, zs, list(reps)

    def monostatic_rcs(self, frequency, theta_deg, phi_deg, mode=0, polarization=1, E_amp=1.0, phase=0.0,
                       excitation=1, integration=1, order=6, want_currents=False):
        """peec_engine_monostatic_rcs. Returns (status, results, currents, report)."""
        lib = load()
        th = np.ascontiguousarray(np.atleast_1d(theta_deg), np.float64)
        ph = np.ascontiguousarray(np.broadcast_to(np.atleast_1d(phi_deg), th.shape), np.float64)
        na = th.size
        o = PeecRcsOptions()
        lib.peec_rcs_options_default(C.byref(o))
        o.polarization, o.E_amp, o.phase, o.excitation = polarization, E_amp, phase, excitation
        o.rcs.integration, o.rcs.surface_order = integration, order
        res = (api.RcsResult * na)()
        I = np.zeros(na * self.ne, np.complex128) if want_currents else None
        rep = PeecPointReport()
        st = lib.peec_engine_monostatic_rcs(self.handle, float(frequency), int(mode), C.byref(o),
                                            th.ctypes.data_as(_dp), ph.ctypes.data_as(_dp), na, res,
                                            I.ctypes.data_as(_cp) if I is not None else None, C.byref(rep))
        return st, list(res), (I.reshape(na, self.ne) if I is not None else None), rep

    def port_sweep(self, frequencies, port, mode=0):
        """peec_engine_port_sweep: (status, z_in, reports)."""
        f = np.ascontiguousarray(frequencies, np.float64)
        z = np.zeros(f.size, np.complex128)
        reps = (PeecPointReport * f.size)()
        p = np.ascontiguousarray(port, np.int32)
        st = load().peec_engine_port_sweep(self.handle, f.ctypes.data_as(_dp), f.size, int(mode), p.ctypes.data_as(_ip),
                                           z.ctypes.data_as(_cp), reps)
        return st, z, list(reps)

    def lightning_response(self, source, n_freq, df, mode=0):
        """peec_engine_lightning_response: (status, time, current, voltage, z_in, reports)."""
        t, i, v = np.zeros(2 * n_freq), np.zeros(2 * n_freq), np.zeros(2 * n_freq)
        z = np.zeros(n_freq, np.complex128)
        reps = (PeecPointReport * n_freq)()
        st = load().peec_engine_lightning_response(self.handle, C.byref(source), n_freq, df, int(mode),
                                                   t.ctypes.data_as(_dp), i.ctypes.data_as(_dp),
                                                   v.ctypes.data_as(_dp), z.ctypes.data_as(_cp), reps)
        return st, t, i, v, z, list(reps)

    def set_profiling(self, on=True):
        load().peec_engine_set_profiling(self.handle, int(on))

    def kernel_stats(self, name):
        st = PeecKernelStats()
        load().peec_engine_kernel_stats(self.handle, KERNELS[name], C.byref(st))
        return {"launches": st.launches, "ms": st.total_ms, "flops": st.flops, "bytes": st.bytes}

    @property
    def far_evaluations(self):
        return load().peec_engine_far_evaluations(self.handle)

    def static_stats(self):
        """Evaluation counts and kernel times of the static assembly, with its
        work in flop-model terms."""
        st = PeecStaticStats()
        if load().peec_engine_static_stats(self.handle, C.byref(st)) != 0:
            raise RuntimeError("peec_engine_static_stats failed")
        flops = (FLOPS_PER_INV_R * (st.far_evaluations + st.near_pair_evaluations)
                 + FLOPS_PER_TRIANGLE_POTENTIAL * st.near_potential_evaluations)
        return {"far_evaluations": st.far_evaluations, "near_potential_evaluations": st.near_potential_evaluations,
                "near_pair_evaluations": st.near_pair_evaluations, "far_ms": st.far_ms, "near_ms": st.near_ms,
                "far_flops": FLOPS_PER_INV_R * st.far_evaluations,
                "near_flops": FLOPS_PER_INV_R * st.near_pair_evaluations
                + FLOPS_PER_TRIANGLE_POTENTIAL * st.near_potential_evaluations, "flops": flops}

    @property
    def smooth_evaluations(self):
        return load().peec_engine_smooth_evaluations(self.handle)


def export_dense_mtx(path, M):
    M = np.ascontiguousarray(M, np.float64)
    if load().matrix_export_dense_mtx(path.encode(), M.ctypes.data_as(_dp), M.shape[0], M.shape[1]) != 0:
        raise RuntimeError("matrix_export_dense_mtx failed")


def import_dense_mtx(path):
    """matrix_import_dense_mtx -> (rows x cols) array."""
    lib = load()
    p = _dp()
    r, c = C.c_size_t(), C.c_size_t()
    if lib.matrix_import_dense_mtx(path.encode(), C.byref(p), C.byref(r), C.byref(c)) != 0:
        raise RuntimeError("matrix_import_dense_mtx failed")
    out = np.ctypeslib.as_array(p, shape=(r.value * c.value,)).copy().reshape(r.value, c.value)
    C.CDLL(None).free(p)
    return out


class Server:
    """Request loop over device-resident engines (include/peec_serve.h):
    admission by HBM budget, factor reuse keyed by (mesh, frequency, mode),
    up to `slots` requests running at once on their own streams, queued
    mixed requests of one mesh batched into shared factor launches (up to
    `max_batch`)."""

    FP64 = 0
    MIXED = 1

    def __init__(self, slots=4, hbm_budget_bytes=0, retarded=False, max_batch=None, **engine_opts):
        lib = load()
        o = PeecServerOptions()
        lib.peec_server_options_default(C.byref(o))
        o.slots = slots
        o.hbm_budget_bytes = int(hbm_budget_bytes)
        if max_batch is not None:
            o.max_batch = int(max_batch)
        o.engine.retarded = int(bool(retarded))
        for k, v in engine_opts.items():
            setattr(o.engine, k, v)
        h = C.c_void_p()
        if lib.peec_server_create(C.byref(o), C.byref(h)) != 0:
            raise RuntimeError("peec_server_create failed")
        self.handle = h
        self._models = []  # the meshes must outlive the server
        self._nports = {}

    @staticmethod
    def engine_bytes(n_edges, n_nodes, retarded):
        """Estimated device bytes of one engine (the unit of admission)."""
        return load().peec_server_engine_bytes(n_edges, n_nodes, int(bool(retarded)))

    def add_mesh(self, model: Model) -> int:
        mid = C.c_int()
        if load().peec_server_add_mesh(self.handle, C.byref(model.mesh), C.byref(model.dual), C.byref(mid)) != 0:
            raise RuntimeError("peec_server_add_mesh failed")
        self._models.append(model)
        return mid.value

    def submit(self, mesh_id, frequency, ports, mode=MIXED) -> int:
        p = np.ascontiguousarray(np.asarray(ports, dtype=np.int32).reshape(-1))
        t = C.c_longlong()
        if load().peec_server_submit(self.handle, mesh_id, frequency, mode, p.ctypes.data_as(_ip), p.size // 2,
                                     C.byref(t)) != 0:
            raise RuntimeError("peec_server_submit rejected the request")
        self._nports[t.value] = p.size // 2
        return t.value

    def wait(self, ticket):
        """(status, z_in array, PeecRequestResult) of a submitted request."""
        n = self._nports.pop(ticket)
        z = np.zeros(n, dtype=np.complex128)
        r = PeecRequestResult()
        st = load().peec_server_wait(self.handle, ticket, z.ctypes.data_as(_cp), C.byref(r))
        return st, z, r

    def memory(self):
        n = C.c_int()
        b = load().peec_server_memory(self.handle, C.byref(n))
        return b, n.value

    def close(self):
        if self.handle:
            load().peec_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
