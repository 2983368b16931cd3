"""MI355X-native (gfx950) multiprecision linear-algebra core with the sdpvibe module API.

Modules mirror the reference package: limbops, rns, mpmat, mpf (more to come). Device work runs
in libsdpvibe_hip.so (hand-written HIP kernels behind a C ABI, include/sdpvibe_hip.h); build it
with `python -m sdpvibe_amd.build`.
"""
__all__ = ["limbops", "rns", "mpmat", "mpf"]
