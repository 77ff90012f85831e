"""MI355X (gfx950) swarm log-likelihood evaluation for FEREBUS-v7 models.

Python host mirror of the C ABI in include/ferebus_hip.h.  The arithmetic
runs in libferebus_hip.so (hand-written HIP for gfx950); this package builds
inputs (kernel programs, mean vectors) and calls it.

Imported before torch or anything else that loads the HIP runtime, the
package also asks the runtime for kernel arguments in device memory
(HIP_FORCE_DEV_KERNARG=1, unless the environment already sets it): the
factorisation's dependent launches then start sooner.  Imported later, it
may have no effect.
"""
import os as _os

_os.environ.setdefault("HIP_FORCE_DEV_KERNARG", "1")

from .kernel_expr import (KernelConfig, KernelExpressionError, KernelProgram, compile_kernel,
                          parse)
from .mean import MeanFit, fit_mean
from .gpr import MultiSwarm, SwarmLikelihood
from ._abi import FbError, device_name, library_path, load
from . import lbfgs
from . import pso

__all__ = ["KernelConfig", "KernelExpressionError", "KernelProgram", "compile_kernel", "parse",
           "MeanFit", "fit_mean", "SwarmLikelihood", "MultiSwarm", "FbError", "device_name", "library_path", "load", "lbfgs", "pso"]
