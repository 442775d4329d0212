"""umpalumpa_amd: MI355X-native Fourier-reconstruction insertion (and the Fourier-processing
crop that feeds it) behind the reference's AFR/AFP operation API.

The compute path is the C ABI in include/umpalumpa_amd.h, implemented by hand-written gfx950
HIP kernels in umpalumpa_amd/csrc and loaded from umpalumpa_amd/libumpalumpa_amd.so.
"""
from ._lib import load, UmpaError  # noqa: F401

__version__ = "0.1.0"
