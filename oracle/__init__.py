"""CPU oracle for the 8th-order stencil path (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg may
import this package; it checks results and is never the thing measured or shipped.

* ``Oracle``     -- ctypes binding of liboracle.so, the C restatement of the
                    reference CPU path (oracle/stencil_oracle.c).
* ``RefLib``     -- ctypes binding of oracle/_ref/libref_*.so, the reference's
                    own CPU path compiled from its sources (oracle/Makefile).
"""
from .lib import Oracle, RefLib, ORACLE_DIR, load_oracle, load_ref, ref_available

__all__ = ["Oracle", "RefLib", "ORACLE_DIR", "load_oracle", "load_ref", "ref_available"]
