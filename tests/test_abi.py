"""CPU checks of the boundary: the library loads, exports every symbol that
include/jfem.h declares, and the host mirror's error paths behave like the
reference's.  No compute calls (there is no GPU here)."""
import ctypes
import os
import re
import warnings

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "jfem.h")).read()
    return sorted(set(re.findall(r"\b(jfm_[a-z_]+)\s*\(", text)))


def test_header_and_binding_agree():
    from jfem import _lib
    assert sorted(_lib.SYMBOLS) == declared_symbols()


def test_library_exports_all_symbols():
    from jfem import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libjfem_hip.so not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name
    lib = _lib.load()
    assert lib.jfm_learned_blob_bytes() == 24 + 2 * (64 * 24 + 64 * 64 + 6 * 64) + 4 * (64 + 64 + 6 + 48 + 1)


def test_unsupported_element_raises_before_gpu():
    """The reference's only elasticity test: a Seg3 in a 3-D continuum problem
    throws (test/test_problems_elasticity_assemble_3d_seg3.jl:8-12)."""
    import jfem
    p = jfem.Problem(jfem.Elasticity, "body", 3)
    p.set_geometry(np.zeros((3, 3)))
    p.add_elements("Seg3", [[0, 1, 2]])
    with pytest.raises(RuntimeError, match="unsupported elements of type Seg3"):
        jfem.assemble(p, 0.0)


def test_empty_problem_and_nonempty_assembly_warn():
    import jfem
    p = jfem.Problem(jfem.Elasticity, "body", 3)
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        assert jfem.assemble(p, 0.0) is None
    assert any("no elements" in str(x.message) for x in w)
    p.assembly.f = np.zeros(3)
    p.add_elements("Tet10", np.zeros((1, 10), np.int64))
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        jfem.assemble(p, 0.0)
    assert any("not empty" in str(x.message) for x in w)


def test_dirichlet_collocation_host_side():
    import jfem
    bc = jfem.Problem(jfem.Dirichlet, "fixed", 3, "displacement")
    bc.add_elements("Poi1", [4, 7], displacement_1=0.0, displacement_3=0.5)
    jfem.assemble(bc, 0.0)
    rows, vals = bc.assembly.fixed
    assert sorted(rows.tolist()) == [12, 14, 21, 23]
    assert dict(zip(rows.tolist(), vals.tolist()))[14] == 0.5


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from jfem import _lib
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(ImportError, match="not built"):
        _lib.load()
