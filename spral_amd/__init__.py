"""MI355X (gfx950) numeric factorize-and-solve backend for SPRAL SSIDS.

spral_amd.ssids mirrors the SSIDS API; the numeric work runs in
libspral_ssids_hip.so (C-ABI declared in include/spral_ssids_hip.h).
"""
from . import ssids  # noqa: F401
