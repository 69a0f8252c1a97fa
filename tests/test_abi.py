"""The C-ABI library loads and exports every function include/graspa_engine.h
declares (no GPU work: symbol lookup only)."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "graspa_engine.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(graspa_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("graspa_box_create", "graspa_site_energies", "graspa_trial_energies", "graspa_single_body_delta",
                 "graspa_ewald_delta", "graspa_ewald_accept", "graspa_commit_move", "graspa_commit_insert",
                 "graspa_commit_delete", "graspa_total_energy", "graspa_widom_batch", "graspa_mc_cycles"):
        assert must in names


def test_library_exports_all_declared_symbols():
    import ctypes
    from graspa_amd import engine
    L = ctypes.CDLL(engine.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from graspa_amd import engine
    assert set(declared_functions()) <= set(engine._SIGS)
