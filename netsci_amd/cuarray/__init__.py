"""cuarray — CuArray host/device buffers (reference: cuarray/python/cuarray/__init__.py:1-2,
cuarray/swig/python_cuarray*.i)."""
from .._netsci import FloatCuArray, IntCuArray, FloatCuArrayRow, IntCuArrayRow  # noqa: F401

__all__ = ["FloatCuArray", "IntCuArray", "FloatCuArrayRow", "IntCuArrayRow"]
