"""MI355X Monte Carlo / RQMC pricing engine: Python harness over the C ABI
declared in include/pe.h (the kernels and host plan compiler live in
csrc/ and are built into _lib/librpl_amd.so)."""
from . import abi, request, wire  # noqa: F401
