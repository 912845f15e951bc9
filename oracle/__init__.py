"""CPU restatement of the WaveFunctionSimPython propagation path (test infrastructure).

This package is the checker, not the product.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product package ``wavefunctionsimpython_amd`` never imports it
and fails loudly when its HIP library is missing.
"""

from .wfs_oracle import *  # noqa: F401,F403
from . import fno_oracle  # noqa: F401,E402
