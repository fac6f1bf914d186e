"""MI355X-native (gfx950) SPH per-timestep hot path of reyn.

The compute lives in ``libreyn_amd.so`` (hand-written HIP behind the C-ABI in
``include/reyn_amd.h``); ``capi`` binds it, ``device`` wraps contexts and
device buffers, ``sim`` mirrors reyn's host API, ``scenes`` generates the
seeded synthetic scenes.
"""

__all__ = ["capi", "device", "sim", "scenes"]
