"""The RHS callable type (numerics/time_integration/rhs_type.py:22-29)."""

from typing import Any, Callable, Dict

FlowFieldMap = Dict[str, Any]
RHS = Callable[[Any, FlowFieldMap, FlowFieldMap], FlowFieldMap]
