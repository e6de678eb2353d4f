"""clmsoil: MI355X (gfx950) implicit soil-column step of ClimaLand.jl.

The compute path is libclmsoil.so (HIP, C ABI in include/clm_soil.h);
this package is its host-side mirror of the reference model interface
(clmsoil.model) plus the synthetic workload generator (clmsoil.inputs).
"""
from . import _lib, inputs  # noqa: F401

__all__ = ["_lib", "inputs", "model"]
