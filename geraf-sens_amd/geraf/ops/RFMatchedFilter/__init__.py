"""Matched-filter ops; lazy exports as in the reference (RFMatchedFilter/__init__.py:6-18)."""
from importlib import import_module

_EXPORTS = {
    "matched_filter_cuda": ".matched_filter_cuda",
    "faster_matched_filter_cuda": ".faster_matched_filter_cuda",
}

__all__ = list(_EXPORTS)


def __getattr__(name):
    if name not in _EXPORTS:
        raise AttributeError(name)
    return getattr(import_module(_EXPORTS[name], __name__), name)
