"""Host code of the drop-in layer (corr_amd/dropin/corr_dropin.cpp) and the
check program (tests/dropin_check.cpp) built with AddressSanitizer and
UndefinedBehaviorSanitizer and run on the CPU: argument errors, a truncated
input file, and a full input read up to the first device call (which, without
a GPU, must fail with a status, not a crash).  The GPU run of the same build
through the whole sequence is in test_gpu_dropin.py."""
import os
import subprocess

import numpy as np
import pytest
import torch

from oracle import posthf as O
from tests import _harness
from tests import _synth as S

SAN_ENV = {"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0", "UBSAN_OPTIONS": "print_stacktrace=1"}


def _san_clean(res):
    err = res.stderr
    assert "AddressSanitizer" not in err and "runtime error" not in err and "LeakSanitizer" not in err, err[-3000:]


def write_input(path):
    from tests.test_gpu_dropin import write_input as w
    N, nocc_sp, naux = 8, 2, 24
    h_ao, eri = S.model_hamiltonian(N, seed=31)
    C, eps, F = S.rhf_orthonormal(h_ao, eri, nocc_sp)
    val, idx, sc = O.sorted_unique_list(eri)
    J = S.metric(naux, seed=5, cond=1e4)
    X = np.random.default_rng(6).standard_normal((naux, N, N)) * 0.1
    w(path, N, 2 * nocc_sp, naux, 9.1875, C, F, h_ao, val, idx, sc, J, X + X.transpose(0, 2, 1))


@pytest.fixture(scope="module")
def san_bin():
    try:
        return _harness.dropin_check(sanitize=True)
    except RuntimeError as exc:
        pytest.skip(str(exc))


def run(binary, *args):
    return subprocess.run([str(binary), *map(str, args)], capture_output=True, text=True, timeout=120,
                          env={**os.environ, **SAN_ENV})


def test_usage_errors(san_bin):
    for args in ((), ("x", "--bogus")):
        res = run(san_bin, *args)
        assert res.returncode == 1 and "usage" in res.stderr
        _san_clean(res)


def test_truncated_input(san_bin, tmp_path):
    full = tmp_path / "full.bin"
    write_input(full)
    data = full.read_bytes()
    for cut in (0, 40, len(data) // 2, len(data) - 8):
        p = tmp_path / f"cut{cut}.bin"
        p.write_bytes(data[:cut])
        res = run(san_bin, p)
        assert res.returncode == 2 and "short input file" in res.stderr, (cut, res.stderr)
        _san_clean(res)


@pytest.mark.skipif(torch.cuda.is_available(), reason="with a GPU the whole sequence runs (test_gpu_dropin.py)")
def test_full_input_without_device(san_bin, tmp_path):
    inp = tmp_path / "sys.bin"
    write_input(inp)
    res = run(san_bin, inp)
    # every array is read; the first device call reports CORR_ENODEV and the process exits with status 3
    assert res.returncode == 3, res.stderr
    assert "corr_ctx_create failed with status 6" in res.stderr
    _san_clean(res)
