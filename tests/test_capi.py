"""The C-ABI library loads and exports every declared entry point (CPU)."""
import re
from pathlib import Path

HEADER = Path(__file__).resolve().parents[1] / "include" / "spral_ssids_hip.h"


def declared():
    txt = HEADER.read_text()
    return sorted(set(re.findall(r"\b(spral_ssids_hip_\w+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared()
    for must in ["spral_ssids_hip_create_symbolic", "spral_ssids_hip_factor",
                 "spral_ssids_hip_solve_fwd", "spral_ssids_hip_solve_diag",
                 "spral_ssids_hip_solve_diag_bwd", "spral_ssids_hip_solve_bwd",
                 "spral_ssids_hip_enquire_indef", "spral_ssids_hip_alter",
                 "spral_ssids_hip_get_contrib", "spral_ssids_hip_destroy_numeric"]:
        assert must in names


def test_library_exports_all_declared(lib):
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_default_options(lib):
    from spral_amd import _lib
    o = _lib.default_options()
    # src/ssids/datatypes.f90:188-284; u and multiplier are default-real
    # literals there, single precision widened to double
    import numpy as np
    assert o.u == float(np.float32(0.01)) and o.small == 1e-20 and o.pivot_method == 2
    assert o.failed_pivot_method == 1 and o.action == 1
    assert o.multiplier == float(np.float32(1.1))
