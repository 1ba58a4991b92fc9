"""Signal-tracer ops (the reference's RFSignalTracer/__init__.py:6-18 exports) plus
the fused lensless renderer.  Imported eagerly, as in RFMatchedFilter."""
from .bilinear_raysample_cuda import bilinear_raysample_cuda
from .lensless_tracer import lensless_signal_tracer
from .signal_tracer_cuda import signal_tracer_cuda

__all__ = ["signal_tracer_cuda", "bilinear_raysample_cuda", "lensless_signal_tracer"]
