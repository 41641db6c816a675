"""hgdt -- MI355X-native global assembly and Jacobi-PCG solve for dune-gdt's diffusion path.

The compute path is libhgdt.so (HIP kernels for gfx950 behind the C ABI of include/hgdt.h);
this package is its Python face with the names of dune.gdt.
"""
from . import _lib, fields
from ._lib import (HgdtError, device_error, integrand_error, linear_solver_failed, mapper_error, not_supported,
                   operator_error, out_of_memory, shapes_do_not_match, space_error, wrong_input_given)
from .gdt import *  # noqa: F401,F403
from .gdt import __all__ as _gdt_all

__all__ = list(_gdt_all) + ["HgdtError", "load_library", "library_path"]


def load_library():
    """Load libhgdt.so; raises ImportError when it is not built."""
    return _lib.load()


def library_path():
    return _lib.LIB_PATH
