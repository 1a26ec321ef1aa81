"""MI355X-native horizon and lighting hot path of lunarscout."""

from .errors import (  # noqa: F401
    ComputeBackendError,
    ContractValidationError,
    HipError,
    HorizonGenerationError,
    InputError,
    LunarscoutError,
    OperationCancelledError,
    ProductGenerationError,
)
from .geometry import DemGrid, ProjectionParameters  # noqa: F401
