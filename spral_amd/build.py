"""Build libspral_ssids_hip.so in-tree with hipcc for gfx950.

Every .hip and .cpp file under spral_amd/csrc is compiled with hipcc for
gfx950 (host files too: they share __host__ __device__ helpers with the
kernels, and one toolchain produces the whole C-ABI library).  Objects are
cached by source mtime under spral_amd/_build; the link always runs.
"""
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
BUILD = HERE / "_build"
LIB = HERE / "libspral_ssids_hip.so"
INCLUDE = HERE.parent / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = ["-O3", "-fPIC", "-std=c++17", f"-I{INCLUDE}", f"-I{CSRC}",
          "-Wno-unused-result"]
# Packed-FP32 instructions are kept out of device code on this target.
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-munsafe-fp-atomics"]


def _run(cmd):
    print(" ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True)


def build(force: bool = False, verbose: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    headers = list(CSRC.glob("*.hpp")) + list(INCLUDE.glob("*.h"))
    hdr_mtime = max((h.stat().st_mtime for h in headers), default=0.0)
    objs = []
    procs = []
    for src in sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp")):
        obj = BUILD / (src.name + ".o")
        objs.append(obj)
        if (not force and obj.exists()
                and obj.stat().st_mtime >= max(src.stat().st_mtime, hdr_mtime)):
            continue
        cmd = [HIPCC, *COMMON, *DEVICE, "-x", "hip"]
        cmd += ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(str(c) for c in cmd), flush=True)
        procs.append(subprocess.Popen([str(c) for c in cmd]))
    bad = [p for p in procs if p.wait() != 0]
    if bad:
        raise RuntimeError("hipcc compilation failed")
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs,
          "-o", tmp, "-lamdhip64"])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> None:
    """Compile the CPU oracle (and the reference-analyse shim where the
    reference tree is mounted).  Test infrastructure, never linked by the
    library."""
    _run(["make", "-s", "-C", HERE.parent / "oracle", "all"])


def build_probe() -> None:
    """Compile the test-only MFMA layout probe (tests/hip/libprobe.so)."""
    src = HERE.parent / "tests" / "hip" / "probe.hip"
    if src.exists():
        _run([HIPCC, "-O2", "-fPIC", "-shared", *DEVICE, f"-I{CSRC}", "-x", "hip", src,
              "-o", src.with_name("libprobe.so")])


if __name__ == "__main__":
    build(force="--force" in sys.argv)
    build_oracle()
