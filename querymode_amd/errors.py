"""QueryModeError with the reference's error codes (src/errors.ts:7-32)."""
from __future__ import annotations

# C-ABI status -> QueryModeError code (include/qmgpu.h)
STATUS_CODES = {
    -1: "TABLE_NOT_FOUND",
    -2: "COLUMN_NOT_FOUND",
    -3: "INVALID_FORMAT",
    -4: "SCHEMA_MISMATCH",
    -5: "INVALID_FILTER",
    -6: "INVALID_AGGREGATE",
    -7: "MEMORY_EXCEEDED",
    -8: "QUERY_FAILED",
    -9: "UNSUPPORTED",
    -10: "INVALID_ARGUMENT",
    -11: "OVERFLOW",
}


class QueryModeError(Exception):
    """Error carrying a machine-readable ``code`` (same names as the reference)."""

    def __init__(self, code: str, message: str):
        super().__init__(message)
        self.code = code
        self.message = message

    def __repr__(self) -> str:  # pragma: no cover - cosmetic
        return f"QueryModeError({self.code!r}, {self.message!r})"
