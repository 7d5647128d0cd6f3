"""The C-ABI library: loads, exports every symbol include/*.h declares, and
its host-side geometry producers match the oracle (no GPU needed)."""
from __future__ import annotations

import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

from peeeeeeeeeeeeeeeeeeec_amd import ctypes_api as api
from peeeeeeeeeeeeeeeeeeec_amd import lib as plib
from peeeeeeeeeeeeeeeeeeec_amd import meshgen
from tests import checker as ck
from tests.conftest import gpu_available

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    """Function names of every prototype (a line starting with a return type)."""
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        for m in re.finditer(r"^(?!static|typedef|#)[A-Za-z_][A-Za-z0-9_ \t\*]*?\b([A-Za-z_][A-Za-z0-9_]*)\s*\(",
                             text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = plib.load()
    declared = declared_functions()
    assert len(declared) > 60
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(plib.EXPORTED) == declared


def test_struct_sizes_match_reference_layout():
    """sizeof() of the reference's structs as gcc lays them out on x86-64
    from the reference headers (Mesh, QuadPatch, DualMesh, InductanceOptions,
    PotentialOptions, ComplexLUFactorization, HarmonicSystem,
    LUFactorization); the ctypes mirrors and include/*.h must agree."""
    expected = {"Mesh": 120, "QuadPatch": 104, "DualMesh": 64, "InductanceOptions": 40,
                "PotentialOptions": 32, "ComplexLUFactorization": 64, "HarmonicSystem": 136,
                "LUFactorization": 48, "IncidentField": 200, "ScatteringOptions": 8, "RcsOptions": 12,
                "RcsResult": 152, "HarmonicSolution": 32}
    for name, size in expected.items():
        assert C.sizeof(getattr(api, name)) == size, name


SHARED_STRUCTS = ["Mesh", "QuadPatch", "PatchList", "DualMesh", "Complex", "ComplexMatrix", "Matrix",
                  "InductanceOptions", "InductanceMatrix", "PotentialOptions", "PotentialMatrix", "IncidenceMatrix",
                  "ResistanceOptions", "ResistanceMatrix", "LUOptions", "LUFactorization", "ComplexLUOptions",
                  "ComplexLUFactorization", "HarmonicSystem", "HarmonicExcitation", "HarmonicSolution",
                  "HarmonicResidual", "IncidentField", "ScatteringOptions", "RcsOptions", "RcsResult",
                  "LightningSource", "InternalCircuitShunt", "TransientSystem"]
PROBE_HEADERS = ["mesh.h", "dual_mesh.h", "quadrature.h", "complex_matrix.h", "matrix.h", "inductance.h",
                 "potential.h", "incidence.h", "resistance.h", "lu.h", "complex_lu.h", "harmonic.h",
                 "incident_field.h", "scattering.h", "rcs.h", "lightning.h", "internal_circuit.h", "transient.h"]


def layout_probe(tmp_path, include_dir):
    """sizeof and the offset of every field (field names from the ctypes
    mirrors) of the shared structs, as gcc lays them out from the headers in
    include_dir."""
    import subprocess
    lines = ["#include <stdio.h>", "#include <stddef.h>"] + [f'#include "{h}"' for h in PROBE_HEADERS]
    lines.append("int main(void){")
    for name in SHARED_STRUCTS:
        lines.append(f'printf("{name} %zu\\n", sizeof({name}));')
        for field, *_ in getattr(api, name)._fields_:
            lines.append(f'printf("{name}.{field} %zu\\n", offsetof({name}, {field}));')
    lines.append("return 0;}")
    tag = "ref" if include_dir != os.path.join(ROOT, "include") else "ours"
    src = tmp_path / f"layout_{tag}.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / f"layout_{tag}"
    subprocess.run(["gcc", "-std=c17", "-w", "-I", include_dir, str(src), "-o", str(exe), "-lm"], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    return dict(line.split() for line in out if line)


def test_headers_layout_matches_ctypes_mirrors(tmp_path):
    """include/*.h and the ctypes mirrors agree on every size and field offset."""
    got = layout_probe(tmp_path, os.path.join(ROOT, "include"))
    for name in SHARED_STRUCTS:
        cls = getattr(api, name)
        assert int(got[name]) == C.sizeof(cls), name
        for field, *_ in cls._fields_:
            assert int(got[f"{name}.{field}"]) == getattr(cls, field).offset, (name, field)


REF_INCLUDE = "/root/reference/include"


@pytest.mark.skipif(not os.path.isdir(REF_INCLUDE), reason="reference headers not present")
def test_headers_layout_matches_reference_headers(tmp_path):
    """The same probe compiled against the reference's own headers
    (include/*.h of the reference) gives the same size and field offsets for
    every shared struct: a drop-in caller compiled against either header set
    sees one layout."""
    ours = layout_probe(tmp_path, os.path.join(ROOT, "include"))
    ref = layout_probe(tmp_path, REF_INCLUDE)
    assert ours == ref


def test_headers_compile_to_reference_layout(tmp_path):
    src = tmp_path / "sz.c"
    src.write_text(
        '#include <stdio.h>\n#include "harmonic.h"\n#include "lu.h"\n#include "dual_mesh.h"\n'
        'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu %zu", sizeof(Mesh), sizeof(QuadPatch),'
        'sizeof(DualMesh), sizeof(InductanceOptions), sizeof(PotentialOptions),'
        'sizeof(ComplexLUFactorization), sizeof(HarmonicSystem), sizeof(LUFactorization));return 0;}')
    exe = tmp_path / "sz"
    import subprocess
    subprocess.run(["gcc", "-std=c17", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    assert out == ["120", "104", "64", "40", "32", "64", "136", "48"]


def test_rcs_headers_compile_to_reference_layout(tmp_path):
    """incident_field.h / scattering.h / rcs.h: sizes and two field offsets as
    gcc lays out the reference's headers (200, 8, 12, 152; E0 at 120,
    pulse_beta at 192, sigma at 136)."""
    src = tmp_path / "sz.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "rcs.h"\n#include "scattering.h"\n'
        'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu", sizeof(IncidentField), sizeof(ScatteringOptions),'
        'sizeof(RcsOptions), sizeof(RcsResult), offsetof(IncidentField, E0), offsetof(IncidentField, pulse_beta),'
        'offsetof(RcsResult, sigma));return 0;}')
    exe = tmp_path / "sz"
    import subprocess
    subprocess.run(["gcc", "-std=c17", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    assert out == ["200", "8", "12", "152", "120", "192", "136"]


@pytest.mark.parametrize("which", ["plate_msh", "box"])
def test_host_geometry_matches_oracle(which):
    if which == "plate_msh":
        m = plib.Model(msh=ck.PLATE_MSH)
        o = ck.OracleModel(msh=ck.PLATE_MSH)
    else:
        xyz, quads = meshgen.box(1.0, 0.6, 0.4, 5, 3, 2)
        m = plib.Model(xyz, quads)
        o = ck.OracleModel(xyz, quads)
    try:
        a, b = api.mesh_arrays(m.mesh), api.mesh_arrays(o.mesh)
        for k in a:
            assert np.array_equal(a[k], b[k]), k
        da, db = api.dual_arrays(m.dual), api.dual_arrays(o.dual)
        assert np.array_equal(da["node_region_areas"], db["node_region_areas"])
        assert np.array_equal(da["edge_region_areas"], db["edge_region_areas"])
        for x, y in zip(da["node_patches"] + da["edge_patches"], db["node_patches"] + db["edge_patches"]):
            assert np.array_equal(x, y)
    finally:
        m.close()
        o.close()


def test_host_gauss_legendre_matches_oracle():
    lib = plib.load()
    for order in (2, 3, 4, 5, 6, 8, 10, 12, 16):
        x1, w1, x2, w2 = ((C.c_double * 16)() for _ in range(4))
        assert lib.gauss_legendre_rule(order, x1, w1) == 0
        ck.oracle().oracle_gauss_legendre(order, x2, w2)
        assert list(x1[:order]) == list(x2[:order]) and list(w1[:order]) == list(w2[:order])
    assert lib.gauss_legendre_rule(7, (C.c_double * 16)(), (C.c_double * 16)()) == -1


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU behaviour")
def test_gpu_entry_points_fail_loudly_without_gpu():
    m = plib.Model(*meshgen.plate(2, 2))
    try:
        with pytest.raises(RuntimeError):
            m.potential()
        st, _, _ = plib.complex_lu_factor_solve(np.eye(3, dtype=complex), np.ones(3))
        assert st != 0
    finally:
        m.close()


def test_kernel_kinds_match_header():
    """The Python names of the per-kernel timers map to the PeecKernel values
    include/peec_gpu.h declares, one for one."""
    text = open(os.path.join(ROOT, "include", "peec_gpu.h")).read()
    declared = {m.group(1).lower(): int(m.group(2)) for m in re.finditer(r"PEEC_KERNEL_([A-Z0-9_]+)\s*=\s*(\d+)", text)}
    assert declared == plib.KERNELS
