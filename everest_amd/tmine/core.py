"""Native core library import (reference: system/tmine/core.py:1-7)."""
from everest_amd.core import (corelib, data, GPUWorker, GPUWorkerLibManager, SingleGPUExecution,  # noqa: F401
                              SingleGPUExecutionDiv, SingleGPURunBatch, SingleGPUExecutionPause, allCount,
                              MultiGPUExecutionNaive, MultiGPUExecutionDyn, MultiGPUExecutionDynPause)
