"""CPU oracle for the Horn-Schunck hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the benchmark's ``cpu_baseline``
leg import this package, and only as the checker / CPU baseline.  The product
(``hindemith_amd`` + ``libhmx.so``) never imports it and has no CPU fallback.

Two layers:

* ``liboracle.so`` (``hm_oracle.c``): C restatement of the reference kernels
  (fp32 in the reference's rounding order, and fp64), OpenMP over rows.
* numpy restatements below: written independently of the C code, used to pin
  it (``np_*``).
"""
from .oracle import *  # noqa: F401,F403
