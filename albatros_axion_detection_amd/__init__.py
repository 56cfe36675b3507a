"""MI355X-native ALBATROS baseband hot path.

4/1-bit unpack -> Wiener-filtered inverse PFB -> forward re-PFB
(upchannelization) -> per-channel cross-antenna correlation, as hand-written
gfx950 HIP kernels in ``libalbatros_hip.so`` behind the call surface of the
reference's Python modules:

=====================================  ========================================
reference module                       here
=====================================  ========================================
src/correlations/unpacking.py          ``unpacking``
src/correlations/correlations.py       ``correlations``
src/correlations/correlations_gpu.py   ``correlations.avg_xcorr_all_ant_gpu``
src/correlations/baseband_data_classes ``baseband_data_classes``
src/utils/pfb_utils.py, pfb_gpu_utils  ``pfb``
src/utils/pycufft.py                   ``pfb.rfft / irfft / fft / ifft``
scripts/xcorr/helper_gpu_stream.py     ``repfb``
=====================================  ========================================

Submodules import torch lazily; the library is loaded on first use and there
is no CPU fallback.
"""

__version__ = "0.1.0"


def library_path():
    from ._lib import LIB_PATH
    return LIB_PATH
