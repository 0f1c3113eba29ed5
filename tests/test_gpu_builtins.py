"""Lane layouts of the compiler builtins the kernels are written against,
checked on the device through tests/hip/libprobe.so (test-only library)."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PROBE = Path(__file__).resolve().parent / "hip" / "libprobe.so"


@pytest.fixture(scope="module")
def probe(lib):
    if not PROBE.exists():
        from spral_amd import build
        build.build_probe()
    h = C.CDLL(str(PROBE))
    h.probe_mfma_f64.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    h.probe_mfma_f32.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    h.probe_wave_max_u64.argtypes = [C.c_void_p, C.c_void_p]
    h.probe_table_find.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
    return h


def test_mfma_f64_16x16x4_layout(probe):
    """A[i][k] in lane i+16k, B[k][j] in lane j+16k, D[i][j] in lane
    j+16(i%4), register i/4 -- the layout k_gather/k_update assume."""
    rng = np.random.default_rng(1)
    A = rng.standard_normal((16, 4))
    B = rng.standard_normal((4, 16))
    out = np.zeros(256)
    assert probe.probe_mfma_f64(A.ctypes.data, B.ctypes.data, out.ctypes.data) == 0
    D = A @ B
    got = np.empty((16, 16))
    for i in range(16):
        for j in range(16):
            got[i, j] = out[(j + 16 * (i % 4)) * 4 + i // 4]
    assert np.allclose(got, D, rtol=1e-14, atol=1e-14)


def test_mfma_f32_16x16x4_layout(probe):
    """Same operand layout as the FP64 instruction, but D[i][j] in lane
    j+16(i/4), register i%4 -- the layout of the mixed-precision gather."""
    rng = np.random.default_rng(2)
    A = rng.standard_normal((16, 4)).astype(np.float32)
    B = rng.standard_normal((4, 16)).astype(np.float32)
    out = np.zeros(256, np.float32)
    assert probe.probe_mfma_f32(A.ctypes.data, B.ctypes.data, out.ctypes.data) == 0
    D = A.astype(np.float64) @ B.astype(np.float64)
    got = np.empty((16, 16))
    for i in range(16):
        for j in range(16):
            got[i, j] = out[(j + 16 * (i // 4)) * 4 + i % 4]
    if not np.allclose(got, D, rtol=1e-5, atol=1e-5):
        # name the layout the hardware has, for the message
        where = {}
        for lane in range(64):
            for r in range(4):
                hit = np.argwhere(np.isclose(D, out[lane * 4 + r], rtol=1e-5, atol=1e-6))
                where[(lane, r)] = [tuple(h) for h in hit[:2]]
        pytest.fail(f"f32 MFMA layout differs: {dict(list(where.items())[:12])}")


@pytest.mark.parametrize("seed", range(6))
def test_dpp_wave_max_u64(probe, seed):
    """The pivot search's wave maximum over 64-bit keys (non-negative
    doubles' bit patterns), including ties and keys that differ only in the
    low word, returns the true maximum in every lane."""
    rng = np.random.default_rng(seed)
    vals = np.abs(rng.standard_normal(64)) * 10.0 ** rng.integers(-5, 5, 64)
    if seed % 2:
        vals[rng.integers(0, 64, 5)] = vals.max()          # ties
    if seed == 4:
        base = vals.max()
        vals[7] = np.nextafter(base, np.inf)                # differs in the low word
    if seed == 5:
        vals[:] = 0.0
    keys = vals.view(np.uint64).copy()
    out = np.zeros(64, np.uint64)
    assert probe.probe_wave_max_u64(keys.ctypes.data, out.ctypes.data) == 0
    assert np.all(out == keys.max())


@pytest.mark.parametrize("n", [1, 5, 64, 65, 3000])
def test_table_find(probe, n):
    """The wave-ballot search that maps a workgroup to its front in a level
    table equals searchsorted, including empty segments and workgroups past
    the total (-1)."""
    rng = np.random.default_rng(n)
    cnt = rng.integers(0, 5, n)
    cnt[rng.random(n) < 0.3] = 0
    off = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int32)
    nb = int(off[-1]) + 3
    out = np.zeros(nb, np.int32)
    assert probe.probe_table_find(off.ctypes.data, n, nb, out.ctypes.data) == 0
    b = np.arange(nb)
    exp = np.where(b < off[-1], np.searchsorted(off, b, side="right") - 1, -1)
    assert np.array_equal(out, exp)
