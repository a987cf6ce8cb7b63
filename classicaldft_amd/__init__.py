"""classicaldft_amd: MI355X-native FMT / DFT hot path of classicalDFT.

Python front end over the C ABI in ``include/cdft.h`` (``libcdft.so``).  The
library is required: importing the bindings raises if it is missing, there is
no CPU fallback.
"""
from .cdft import (  # noqa: F401
    CDFT_OK, CDFT_ETA_TOO_LARGE, CDFT_UPHILL_LIMIT, FP64, MIXED, WEIGHTS_ACCURATE, WEIGHTS_REFERENCE, AO_FORCE_EXACT, AO_FORCE_REFERENCE,
    MODEL_NONE, ROSENFELD, RSLT, ESFMT,
    WHITEBEAR_I, WHITEBEAR_II, DEVICE_PTRS, ACCUMULATE, FMT_ONLY, NO_IDEAL, CdftError, EtaTooLarge, System, Mixture, DeviceArray, Fire,
    DDFT, OpenDDFT, Eigenvalues, Arnoldi, SQUARED_MATRIX, DYNAMIC_MATRIX, dst1_3d, mf_kernel_points, FireBatch, FireGroup, UNet, fft_r2c, fft_c2r, lj_hsd, potential_hsd, potential_vdw, potential_watt,
    POT_LJ, POT_TWF, POT_WHDF, MF_QF, MF_ZERO, MF_LE, MF_QE, MF_LF, MF_GAUSS_E, MF_GAUSS_F,
    FIXED_BACKGROUND, HOMOGENEOUS_BOUNDARY, device_count, set_device, get_device, library_path,
)
from .formats import (  # noqa: F401,E402
    read_density, write_density, read_points_file, read_weight_cache, write_weight_cache, weight_cache_name,
    potential_identifier,
)
