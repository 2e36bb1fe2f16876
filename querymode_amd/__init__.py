"""querymode_amd — QueryMode's columnar operator hot path on MI355X (gfx950).

filter -> hash group-by aggregate -> hash-join probe -> top-K, as HIP kernels
behind a C-ABI (include/qmgpu.h), with the reference's DataFrame / operator /
QMCB surface on the host.
"""
from .client import DataFrame, Device, GpuExecutor, GpuTable, QueryMode, QueryResult
from .errors import QueryModeError
from .operators import GpuQueryOperator, drain_pipeline, materialize_rows
from .plan import AggregateOp, BigInt, FilterOp, JoinDescriptor, QueryDescriptor, encode_plan

__all__ = [
    "QueryMode", "DataFrame", "Device", "GpuExecutor", "GpuTable", "QueryResult", "QueryModeError",
    "GpuQueryOperator", "drain_pipeline", "materialize_rows", "AggregateOp", "BigInt", "FilterOp",
    "JoinDescriptor", "QueryDescriptor", "encode_plan",
]
