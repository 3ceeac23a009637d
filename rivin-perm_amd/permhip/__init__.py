"""permhip — exact modular permanents on AMD MI355X (gfx950).

The GPU work runs in hand-written HIP kernels in ``libpermhip.so`` (C ABI in
``include/permhip.h``), reached through ctypes.  Planning, CRT and
preprocessing stay on the host in exact Python integers.

Modules:
  _native     ctypes binding of the C ABI (fails loudly without a device)
  planner     bounds, CRT prime plans, Graham--Lehmer moduli, CRT
  binary_prep 0/1 matrix normalisation, matching/SCC blocks, bounds
  backends    objects shaped like the reference's native extensions
  batch       many matrices x all their primes in one persistent launch
  serve       continuous-batching request loop
  build       hipcc build recipe
"""

from . import planner  # noqa: F401

__all__ = ["planner"]
