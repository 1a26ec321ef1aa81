"""Error types of the public surface (reference src/lunarscout/errors.py:7-75).

``LunarscoutError`` carries a stable ``code`` and ``details`` like the
reference. ``HipError`` takes the place of the reference's ``CudaError`` for
device-unavailable and device-execution failures.
"""

from __future__ import annotations

from typing import Any, Mapping


class LunarscoutError(Exception):
    def __init__(self, message: str, *, code: str, details: Mapping[str, Any] | None = None):
        super().__init__(message)
        self.code = code
        self.details = dict(details or {})


class InputError(LunarscoutError, ValueError):
    pass


class ComputeBackendError(LunarscoutError, RuntimeError):
    pass


class HipError(ComputeBackendError):
    """A GPU backend failure: ``hip_*_unavailable`` or ``hip_*_execution_failed``."""


class HorizonGenerationError(LunarscoutError, RuntimeError):
    pass


class ProductGenerationError(LunarscoutError, RuntimeError):
    pass


class OperationCancelledError(LunarscoutError):
    pass


class ContractValidationError(ValueError):
    """Raised before any device work when an array contract is invalid."""


class HipBackendError(RuntimeError):
    """The HIP library or a usable device is unavailable (cuda_backend.py:19-20 analogue)."""


class HipExecutionError(RuntimeError):
    """A HIP runtime call or kernel launch failed."""
