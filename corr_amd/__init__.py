"""corr_amd: MI355X-native post-HF contraction path of corr.

The compute lives in libcorr_gpu.so (hand-written gfx950 HIP kernels behind
the C-ABI of include/corr_abi.h); this package is the PyTorch-ROCm front-end.
"""
from ._abi import (CorrError, CorrLibraryError, LIB_PATH, PREC_BF16X3, PREC_FP32, PREC_FP64,  # noqa: F401
                   lib)
from .posthf import (MOtrans, ao_unfold, build_FMo, calculate_ri, dgemm, ri_bia_from_j, ri_form_bia,  # noqa: F401
                     ri_mp2_energy_from_j,
                     ri_metric, rimp2, rimp2_energy, rimp2_gradient, run_canonical_mp2, run_canonical_mp2_ri, run_non_canonical_mp2,
                     transform_ri, RIMP2Energy)

from .ccsd import ccsd_energy, ccsd_update, ccsd_ur, ints_so  # noqa: F401
from .scf import build_Fmat, build_Fmat_list, build_Fmat_ri, build_Pmat, run_scf, run_scf_ri  # noqa: F401
from .localize import localize, localize_PM, localize_boys  # noqa: F401

__version__ = "0.1.0"
