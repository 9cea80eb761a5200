"""libddla_amd -- MI355X-native LibDDLA Cholesky path.

Host side of the drop-in: the reference's Python surface (Handle,
DistributedMatrix, ppotrf / ppotrs / pposv, diag_logdet) over the C ABI of
libddla_mi355x.so, whose kernels are hand-written HIP for gfx950.
"""
from ._capi import (
    LIBRARY_PATH,
    DdlaAllocFailedError,
    DdlaBackendUnavailableError,
    DdlaError,
    DdlaExceptionError,
    DdlaInternalError,
    DdlaInvalidHandleError,
    DdlaInvalidValueError,
    check_status,
    lib,
)
from .backend import Backend, Handle
from .batched import (
    BatchedGraph,
    VBatch,
    pposv_vbatched,
    pposv_vbatched_run,
    pposv_vbatched_tune,
    pposv_vbatched_workspace_size,
    set_vbatched_schedule,
    vbatched_schedule,
)
from .matrix import DistributedMatrix, descinit
from .ops import (
    FACTOR_BF16X3,
    FACTOR_BF16X6,
    FACTOR_FP16X3,
    FACTOR_FP32,
    diag_logdet,
    pdiag_logdet,
    potrf_bottom_right,
    ppotrf,
    ppotrf_bottom_right,
    ppotrf_run,
    ppotrf_workspace_size,
    pposv,
    pposv_mixed,
    pposv_mixed_run,
    pposv_mixed_workspace_size,
    pposv_run,
    pposv_workspace_size,
    ppotrs,
    ppotrs_run,
    ppotrs_workspace_size,
    ptrtrs,
    trsm,
)

__all__ = [
    "LIBRARY_PATH", "lib", "check_status", "Backend", "Handle", "DistributedMatrix", "descinit",
    "ppotrf", "ppotrs", "pposv", "ptrtrs", "trsm", "potrf_bottom_right", "ppotrf_bottom_right", "ppotrf_workspace_size", "ppotrf_run", "ppotrs_workspace_size", "ppotrs_run",
    "pposv_workspace_size", "pposv_run", "pposv_mixed", "pposv_mixed_workspace_size", "pposv_mixed_run",
    "FACTOR_FP32", "FACTOR_BF16X3", "FACTOR_BF16X6", "FACTOR_FP16X3",
    "diag_logdet", "pdiag_logdet", "VBatch", "BatchedGraph", "pposv_vbatched", "pposv_vbatched_run", "pposv_vbatched_tune",
    "pposv_vbatched_workspace_size", "vbatched_schedule", "set_vbatched_schedule", "DdlaError", "DdlaInvalidHandleError",
    "DdlaInvalidValueError", "DdlaAllocFailedError", "DdlaBackendUnavailableError", "DdlaInternalError",
    "DdlaExceptionError",
]
