"""seisfem_amd -- MI355X-native time loop for SeisFEM's explicit elastic M-PML
masslump solver.

  host      setup arrays the reference's elastic_wave() builds (bit-identical)
  gpu       device operators, media and multi-shot runs (HIP kernels, C ABI)
  formats   the reference's par.txt / mesh / velocity / seismogram text formats
  workloads synthetic benchmark configurations

The native library (libseisfem_amd.so) is required; there is no CPU fallback.
"""
from ._native import lib, last_error, NativeError  # noqa: F401

__all__ = ["lib", "last_error", "NativeError"]
