"""CPU FP64 oracle of the wloop measurement path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
package; it is the FP64 reference the HIP path is tested against and never part
of the product.
"""
from .wl_oracle import *  # noqa: F401,F403
