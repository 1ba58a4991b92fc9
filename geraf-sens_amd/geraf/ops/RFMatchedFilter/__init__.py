"""Matched-filter ops (same exports as the reference's RFMatchedFilter/__init__.py:6-18).

Imported eagerly: the ``_C`` shim loads libgeraf_rf.so only on first call, and
eager binding keeps ``matched_filter_cuda`` the function even after its
submodule of the same name has been imported."""
from .faster_matched_filter_cuda import faster_matched_filter_cuda
from .matched_filter_cuda import matched_filter_cuda

__all__ = ["matched_filter_cuda", "faster_matched_filter_cuda"]
