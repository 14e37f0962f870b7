"""MI355X (gfx950) GA generation step: HIP kernels behind the C ABI of include/ne_runtime.h.

Import through __graft_entry__.import_package() (the directory name is not a Python identifier);
the package is registered as ``ne_wordle_amd``.
"""
from . import runtime, synth  # noqa: F401
from .runtime import SlabRuntime, load_library  # noqa: F401
