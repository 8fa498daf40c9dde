"""The reference's post-HF sequence (src/main.cc:130-175) run through the C++
drop-in entry points (libcorr_dropin.so) on the GPU; the printed result lines
are parsed and compared with the CPU oracle (10 printed decimals)."""
import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from oracle import ccsd as CC
from oracle import posthf as O
from tests import _harness
from tests import _synth as S

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]


def write_input(path, N, nroe, naux, ion_rep, C, F, H, val, idx, sc, J, Pmn):
    with open(path, "wb") as f:
        np.array([N, nroe, naux, len(val), *sc], dtype=np.int64).tofile(f)
        np.array([ion_rep], dtype=np.float64).tofile(f)
        for a in (C, F, H, val):
            np.ascontiguousarray(a, dtype=np.float64).tofile(f)
        np.ascontiguousarray(idx, dtype=np.uint16).tofile(f)
        for a in (J, Pmn):
            np.ascontiguousarray(a, dtype=np.float64).tofile(f)


@pytest.mark.parametrize("sanitize", [False, True], ids=["plain", "asan_ubsan"])
def test_dropin_sequence_matches_oracle(tmp_path, sanitize):
    """With sanitize=True the check program and the drop-in layer are built with
    AddressSanitizer and UndefinedBehaviorSanitizer (host code only; the
    kernels in libcorr_gpu.so are not instrumented) and must run clean."""
    import os
    N, nocc_sp, naux = 8, 2, 24
    h_ao, eri = S.model_hamiltonian(N, seed=31)
    C, eps, F = S.rhf_orthonormal(h_ao, eri, nocc_sp)
    val, idx, sc = O.sorted_unique_list(eri)
    J = S.metric(naux, seed=5, cond=1e4)
    X = np.random.default_rng(6).standard_normal((naux, N, N)) * 0.1
    Pmn = X + X.transpose(0, 2, 1)
    ion = 9.1875
    inp = tmp_path / "sys.bin"
    write_input(inp, N, 2 * nocc_sp, naux, ion, C, F, h_ao, val, idx, sc, J, Pmn)
    exe = _harness.dropin_check(sanitize=sanitize)
    print(f"running {exe.name}", flush=True)
    env = dict(os.environ)
    if sanitize:
        # the HIP runtime keeps allocations until exit; leaks of the host code are checked on the CPU
        env.update(ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="print_stacktrace=1")
    res = subprocess.run([str(exe), str(inp)], capture_output=True, text=True, timeout=120, env=env)
    sys.stdout.write(res.stdout[-3000:])
    assert res.returncode == 0, res.stderr
    assert "AddressSanitizer" not in res.stderr and "runtime error" not in res.stderr, res.stderr[-3000:]
    out = res.stdout
    vals = [float(x) for x in re.findall(r"^ +EMP2:\s+(-?\d+\.\d+)", out, re.M)]
    assert len(vals) == 2
    prec = O.motrans(C, eri)
    FMo = O.build_fmo(C, F)
    can = O.canonical_mp2(prec, FMo, nocc_sp)
    M, _ = O.ri_metric_eigen(J)
    B = O.transform_ri(C, O.ri_bpq(M, Pmn), nocc_sp)
    e = np.diag(FMo)
    ri = O.rimp2(B, e[:nocc_sp], e[nocc_sp:])
    assert abs(vals[0] - can["EMP2"]) < 2e-10
    assert abs(vals[1] - ri["EMP2"]) < 2e-10
    scs = [float(x) for x in re.findall(r"^\s*EMP2\(SCS\):\s+(-?\d+\.\d+)", out, re.M)]
    assert abs(scs[0] - can["EMP2_SCS"]) < 2e-10 and abs(scs[1] - ri["EMP2_SCS"]) < 2e-10
    ref = CC.ccsd_ur(prec, C, h_ao, 2 * nocc_sp, ion, maxiter=20)
    # "EMP2: <value>" has no newline in the reference, so iteration 0 is printed
    # glued to it ("...0240E(CCSD)= ..."): match the energy field only
    it = [float(x) for x in re.findall(r"E\(CCSD\)= (-?\d+\.\d+)", out)]
    assert len(it) == 20
    assert np.allclose(it, ref["energies"], atol=2e-10)
    el = float(re.search(r"^El:(-?\d+\.\d+)", out, re.M).group(1))
    assert abs(el - ref["e_hf_check"]) < 2e-10
    emp2_so = float(re.search(r"EMP2: (-?\d+\.\d+)", out).group(1))
    assert abs(emp2_so - ref["emp2"]) < 2e-10


def test_dropin_scf_then_post_hf_matches_oracle(tmp_path):
    """main.cc:111-175 through the drop-ins: run_scf from a core guess, then
    the post-HF sequence on the MOs and Fmat it leaves behind."""
    from oracle import scf as SCF
    N, nocc_sp, naux = 8, 2, 24
    H, Sm, eri, *_ = S.nonorthogonal_system(N, seed=41, nocc=nocc_sp)
    T = SCF.calc_s12(Sm)
    C0, _ = SCF.check_initial_guess(np.zeros((N, N)), H, T)
    val, idx, sc = O.sorted_unique_list(eri)
    J = S.metric(naux, seed=7, cond=1e3)
    X = np.random.default_rng(8).standard_normal((naux, N, N)) * 0.1
    Pmn = X + X.transpose(0, 2, 1)
    ion = 4.5
    inp = tmp_path / "scf.bin"
    write_input(inp, N, 2 * nocc_sp, naux, ion, C0, np.zeros((N, N)), H, val, idx, sc, J, Pmn)
    with open(inp, "ab") as f:
        np.ascontiguousarray(Sm, dtype=np.float64).tofile(f)
    print("running corr_dropin_check --scf", flush=True)
    res = subprocess.run([str(_harness.dropin_check()), str(inp), "--scf"], capture_output=True,
                         text=True, timeout=120)
    sys.stdout.write(res.stdout[-3000:])
    assert res.returncode == 0, res.stderr
    out = res.stdout
    ref = SCF.run_scf(H, Sm, lambda P: SCF.fock_dense(eri, P, H), C0, 2 * nocc_sp, ion_rep=ion)
    its = [float(x) for x in re.findall(r"^ *\d+: +(-?\d+\.\d+)", out, re.M)]
    assert len(its) == ref["iterations"]
    assert np.allclose(its, ref["energies"], atol=2e-10)
    assert ("Starting DIIS" in out) == ref["diis"]
    tot = float(re.search(r"TOTAL E:\s+(-?\d+\.\d+)", out).group(1))
    assert abs(tot - ref["total"]) < 2e-10
    C, F = ref["C"], ref["F"]
    prec = O.motrans(C, eri)
    FMo = O.build_fmo(C, F)
    can = O.canonical_mp2(prec, FMo, nocc_sp)
    M, _ = O.ri_metric_eigen(J)
    B = O.transform_ri(C, O.ri_bpq(M, Pmn), nocc_sp)
    e = np.diag(FMo)
    ri = O.rimp2(B, e[:nocc_sp], e[nocc_sp:])
    vals = [float(x) for x in re.findall(r"^ +EMP2:\s+(-?\d+\.\d+)", out, re.M)]
    assert abs(vals[0] - can["EMP2"]) < 2e-10 and abs(vals[1] - ri["EMP2"]) < 2e-10
    cc = CC.ccsd_ur(prec, C, H, 2 * nocc_sp, ion, maxiter=20)
    it = [float(x) for x in re.findall(r"E\(CCSD\)= (-?\d+\.\d+)", out)]
    assert np.allclose(it, cc["energies"], atol=2e-10)


def test_dropin_scf_fock_and_non_canonical_mp2(tmp_path):
    """build_Pmat/build_Fmat drop-ins rebuild F from the final MOs of run_scf
    ("Fcheck:" trace and sum of squares vs the oracle's Fock matrix), and
    run_non_canonical_mp2 prints the oracle's 20 E(sem-loc)/Rsum lines."""
    from oracle import scf as SCF
    N, nocc_sp, naux = 8, 2, 24
    H, Sm, eri, *_ = S.nonorthogonal_system(N, seed=43, nocc=nocc_sp)
    C0, _ = SCF.check_initial_guess(np.zeros((N, N)), H, SCF.calc_s12(Sm))
    val, idx, sc = O.sorted_unique_list(eri)
    J = S.metric(naux, seed=7, cond=1e3)
    X = np.random.default_rng(8).standard_normal((naux, N, N)) * 0.1
    inp = tmp_path / "scf2.bin"
    write_input(inp, N, 2 * nocc_sp, naux, 1.0, C0, np.zeros((N, N)), H, val, idx, sc, J, X + X.transpose(0, 2, 1))
    with open(inp, "ab") as f:
        np.ascontiguousarray(Sm, dtype=np.float64).tofile(f)
    print("running corr_dropin_check --scf --ncmp2", flush=True)
    res = subprocess.run([str(_harness.dropin_check()), str(inp), "--scf", "--ncmp2"],
                         capture_output=True, text=True, timeout=120)
    sys.stdout.write(res.stdout[-2000:])
    assert res.returncode == 0, res.stderr
    out = res.stdout
    ref = SCF.run_scf(H, Sm, lambda P: SCF.fock_dense(eri, P, H), C0, 2 * nocc_sp, ion_rep=1.0)
    P = SCF.build_pmat(ref["C"], 2 * nocc_sp, None, 0.0)
    F = SCF.fock_dense(eri, P, H)
    tr, ss = (float(x) for x in re.search(r"Fcheck: (\S+) (\S+)", out).groups())
    assert abs(tr - np.trace(F)) < 2e-9 and abs(ss - np.sum(F * F)) < 2e-8
    prec = O.motrans(ref["C"], eri)
    E, Rs, _ = O.non_canonical_mp2(prec, O.build_fmo(ref["C"], ref["F"]), nocc_sp, 20)
    got = [float(x) for x in re.findall(r"E\(sem-loc\):\s+(-?\d+\.\d+)", out)]
    assert len(got) == 20 and np.allclose(got, E, atol=2e-10)
