"""MI355X (gfx950) execution layer for the bespoke per-query OLAP pipelines.

The compute path is libbolap.so (hand-written HIP kernels behind the C ABI
in include/bolap.h).  `native` mirrors that ABI with ctypes and `engine`
wraps it for Python callers; nothing here computes results on the CPU.
"""
__all__ = ["native", "engine"]
