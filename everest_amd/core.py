"""Import of the native "corelib" module from everest_amd/lib.

Mirrors system/tmine/core.py:1-7 of the reference: corelib is imported with
RTLD_GLOBAL so generated worker plugins resolve the GPUWorker symbols they
link against.  Raises ImportError when the native build is missing.
"""
import os
import sys

from .abi import LIB_DIR

if LIB_DIR not in sys.path:
    sys.path.insert(0, LIB_DIR)
_flags = sys.getdlopenflags()
sys.setdlopenflags(os.RTLD_GLOBAL | os.RTLD_LAZY)
try:
    import corelib  # noqa: E402
finally:
    sys.setdlopenflags(_flags)

from corelib import data, GPUWorker, GPUWorkerLibManager, SingleGPUExecution, SingleGPUExecutionDiv  # noqa
from corelib import SingleGPURunBatch, SingleGPUExecutionPause  # noqa
from corelib import allCount, MultiGPUExecutionNaive, MultiGPUExecutionDyn, MultiGPUExecutionDynPause  # noqa
