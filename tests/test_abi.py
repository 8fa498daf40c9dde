"""C-ABI and drop-in boundary checks that need no GPU: every function declared
in include/corr_abi.h is exported by libcorr_gpu.so, the drop-in library
exports the reference's C++ signatures, and error paths report status codes."""
import ctypes
import re
import subprocess
from pathlib import Path

import pytest
import torch

ROOT = Path(__file__).resolve().parents[1]


def declared(header):
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(corr_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_entry_point_is_exported():
    import corr_amd
    lib = corr_amd.lib()
    names = declared("corr_abi.h")
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    from corr_amd import _abi
    assert not _abi.MISSING
    assert set(_abi.SIGNATURES) == set(names)


def test_dropin_exports_reference_signatures():
    so = ROOT / "corr_amd" / "libcorr_dropin.so"
    out = subprocess.run(["nm", "-DC", "--defined-only", str(so)], capture_output=True, text=True, check=True).stdout
    for sig in ["build_FMo(systeminfo*, OEints*, pHF*)",
                "MOtrans(systeminfo*, OEints*, TEints*, pHF*)",
                "run_canonical_mp2(systeminfo*, pHF*)",
                "run_canonical_mp2_ri(systeminfo*, pHF*)",
                "transform_ri(systeminfo*, OEints*, pHF*)",
                "ccsd_ur(systeminfo*, OEints*, pHF*)",
                "get_integral(double*, long long&, long long&, long long&, int&, int&, int&, int&)",
                "calculate_ri_arith(int, int, double const*, double const*, double*)"]:
        assert sig in out, sig
    assert "read_transform_ri(" in out


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-device path")
def test_context_without_device_reports_enodev():
    import corr_amd
    lib = corr_amd.lib()
    h = ctypes.c_void_p()
    st = lib.corr_ctx_create(0, ctypes.byref(h))
    assert st == 6          # CORR_ENODEV
    assert not h.value
    assert lib.corr_ctx_destroy(None) == 0
    assert lib.corr_dgemm(None, 0, 0, 1, 1, 1, 1.0, None, 1, 0, None, 1, 0, 0.0, None, 1, 0, 1, 0) == 1  # EINVAL


def test_front_end_rejects_cpu_tensors():
    import corr_amd
    x = torch.zeros(3, 2, 4, dtype=torch.float64)
    with pytest.raises(ValueError, match="GPU only"):
        corr_amd.rimp2_energy(x, torch.zeros(2, dtype=torch.float64), torch.zeros(4, dtype=torch.float64))
    with pytest.raises(ValueError, match="unknown precision"):
        from corr_amd.posthf import _prec
        _prec("fp16")
