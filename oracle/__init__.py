"""FP64 CPU oracle of the Wavey FDTD path -- test infrastructure only.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this package, and only as the checker. The product (``wavey_amd``)
never imports it.
"""
from .oracle import *  # noqa: F401,F403
