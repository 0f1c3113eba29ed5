"""The Fortran binding module driven by a Fortran host on the GPU:
tests/fortran/hip_subtree_driver.f90 (built by oracle/Makefile `boundary`
where the reference tree is mounted) reaches the HIP library only through
symbolic_subtree_base / numeric_subtree_base.  Checked: the documented 5x5
example solves to 1..5, inertia equals the closed form for a shifted 3D
Laplacian, the backward error meets the bar, enquire's D^{-1} blocks give
the same inertia, and alter with the enquired D^{-1} leaves the solve
bitwise unchanged."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DRIVER = Path(__file__).resolve().parents[1] / "oracle" / "_ref" / "boundary" / \
    "hip_subtree_driver"


@pytest.mark.skipif(not DRIVER.exists(), reason="Fortran driver not built")
def test_fortran_host_through_binding(lib):
    r = subprocess.run([str(DRIVER)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    rows = {ln.split()[1]: ln.split()[2:] for ln in r.stdout.splitlines()
            if ln.startswith("CASE")}
    doc = rows["doc5"]
    n, neg, two, rank, nnodes = (int(v) for v in doc[:5])
    bwd, eneg, same = float(doc[5]), int(doc[6]), doc[7]
    x = np.array([float(v) for v in doc[8:13]])
    assert n == 5 and rank == 5 and eneg == neg and same == "T"
    assert np.abs(x - np.arange(1, 6)).max() <= 1e-13 and bwd <= 1e-14
    lap = rows["lap765"]
    n, neg, two, rank, nnodes = (int(v) for v in lap[:5])
    bwd, eneg, same = float(lap[5]), int(lap[6]), lap[7]
    c = [2 * np.cos(np.pi * np.arange(1, k + 1) / (k + 1)) for k in (7, 6, 5)]
    ev = 4.0 - (c[0][:, None, None] + c[1][None, :, None] + c[2][None, None, :])
    assert n == 210 and rank == 210
    assert neg == int((ev < 0).sum()) == eneg
    assert bwd <= 1e-14 and same == "T"
