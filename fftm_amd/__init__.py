"""fftm_amd: MI355X-native backend for fftm's single-GPU ``ffts`` R2C/C2R path.

The transforms run in hand-written HIP kernels for gfx950 (``csrc/``) behind the C ABI
of ``include/fftm_mi.h``; this package is the host-side mirror of the reference's
``external_wrap`` boundary (``wrap``) and of its ``ffts`` frontend (``ffts``).
"""

from ._lib import FftmMiError, LIB_PATH  # noqa: F401
from .wrap import (LogicError, direction, fft, fft_plan_descriptor, fft_wrap_many,  # noqa: F401
                   runtime_api, transpose_4d, fft_direction_name, fill_random_real)
from .ffts import (ffts, ffts_init_options, strategy_4d_pencil_pencil, strategy_4d_slab_slab,  # noqa: F401
                   strategy_4d_slab_native_xzwy, transpose_backend, transform_strategy_4d)

__version__ = "0.1.0"
