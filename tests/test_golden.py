"""Committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py from the CPU oracle on seeded inputs).

CPU tests: the oracle still reproduces every fixture from its stored inputs.
GPU tests: the HIP path reproduces them through the C-ABI:
  FP64 to 1e-10 Eh (energies) / 1e-12 (tensors), FP32 and BF16x3 RI-MP2 to
  2e-6 relative."""
import os
import sys

import numpy as np
import pytest
import torch

from oracle import ccsd as CC
from oracle import posthf as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(HERE, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda")


def _say(m):
    print(m, flush=True)
    sys.stdout.flush()


# ---- CPU: the oracle against the fixtures ----------------------------------------

def test_oracle_rimp2_fixture():
    g = load("rimp2_naux76_o5_v19.npz")
    e = O.rimp2(g["Bia"], g["eps_occ"], g["eps_vir"])
    assert abs(e["EMP2_OS"] - g["e_os"]) < 1e-12 and abs(e["EMP2_SS"] - g["e_ss"]) < 1e-12
    E, dB, deo, dev_ = O.rimp2_gradient(g["Bia"], g["eps_occ"], g["eps_vir"], 1.0, 1.0)
    assert abs(E - g["grad_E"]) < 1e-12
    assert np.allclose(dB, g["grad_B"], rtol=0, atol=1e-13)
    assert np.allclose(deo, g["grad_eo"], rtol=0, atol=1e-13) and np.allclose(dev_, g["grad_ev"], rtol=0, atol=1e-13)


def test_oracle_fourindex_fixture():
    g = load("fourindex_n6.npz")
    N = g["C"].shape[0]
    dense = O.unfold_sorted(N, g["intval"], g["intnums"].astype(np.uint16), g["sortcount"])
    prec = O.motrans(g["C"], dense)
    assert np.allclose(prec, g["prec_ints"], rtol=0, atol=1e-14)
    e = O.canonical_mp2(prec, g["FMo"], 2)
    assert abs(e["EMP2"] - g["e_mp2"]) < 1e-13


def test_oracle_ccsd_fixture():
    g = load("ccsd_n6_e4.npz")
    r = CC.ccsd_ur(g["prec_ints"], g["C"], g["H"], int(g["nroe"]), float(g["ion_rep"]), maxiter=20)
    assert np.allclose(r["energies"], g["energies"], rtol=0, atol=1e-12)
    assert abs(r["e_hf_check"] - g["e_hf_check"]) < 1e-12


# ---- GPU: the HIP path against the fixtures --------------------------------------

@pytest.fixture(scope="module")
def corr():
    import corr_amd
    return corr_amd


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("fp32", 2e-6), ("bf16x3", 2e-6)])
def test_gpu_rimp2_fixture(corr, prec, tol):
    _say(f"golden rimp2 {prec}")
    g = load("rimp2_naux76_o5_v19.npz")
    e = corr.rimp2_energy(dev(g["Bia"]), dev(g["eps_occ"]), dev(g["eps_vir"]), prec)
    if prec == "fp64":
        assert abs(e["EMP2_OS"] - g["e_os"]) < tol and abs(e["EMP2_SS"] - g["e_ss"]) < tol
    else:
        assert abs(e["EMP2"] - g["e_mp2"]) < tol * abs(g["e_mp2"])


@pytest.mark.gpu
def test_gpu_rimp2_gradient_fixture(corr):
    _say("golden rimp2 gradient")
    g = load("rimp2_naux76_o5_v19.npz")
    E, dB, deo, dev_ = corr.rimp2_gradient(dev(g["Bia"]), dev(g["eps_occ"]), dev(g["eps_vir"]), 1.0, 1.0, "fp64")
    assert abs(E - g["grad_E"]) < 1e-10
    assert np.allclose(dB.cpu().numpy(), g["grad_B"], rtol=0, atol=1e-12)
    assert np.allclose(deo.cpu().numpy(), g["grad_eo"], rtol=0, atol=1e-12)
    assert np.allclose(dev_.cpu().numpy(), g["grad_ev"], rtol=0, atol=1e-12)


@pytest.mark.gpu
def test_gpu_fourindex_fixture(corr):
    _say("golden unfold + MOtrans + canonical MP2")
    g = load("fourindex_n6.npz")
    N = g["C"].shape[0]
    idx = torch.from_numpy(g["intnums"].astype(np.int16)).to("cuda")
    dense = corr.ao_unfold(N, dev(g["intval"]), idx, g["sortcount"])
    prec = corr.MOtrans(dev(g["C"]), dense, "fp64")
    assert np.allclose(prec.cpu().numpy(), g["prec_ints"], rtol=0, atol=1e-12)
    e = corr.run_canonical_mp2(prec, dev(g["FMo"]), 2)
    assert abs(e["EMP2"] - g["e_mp2"]) < 1e-10


@pytest.mark.gpu
def test_gpu_ccsd_fixture(corr):
    _say("golden ccsd 20 iterations")
    g = load("ccsd_n6_e4.npz")
    r = corr.ccsd_ur(dev(g["prec_ints"]), dev(g["C"]), dev(g["H"]), int(g["nroe"]), float(g["ion_rep"]),
                     maxiter=20)
    assert np.allclose(r["energies"], g["energies"], rtol=0, atol=1e-10)
    assert abs(r["emp2"] - float(g["emp2"])) < 1e-10
    assert abs(r["e_hf_check"] - float(g["e_hf_check"])) < 1e-10


def test_oracle_water_fixture():
    """The real-water fixture: the oracle's RI-MP2 on the stored B gives the
    stored energies; the pair-loop and dense forms agree; RHF and MP2 lie in
    the known def2-SVP ranges; no auxiliary function was cut by the metric;
    and n separated copies (shuffled indices) give exactly n times the
    energies (size extensivity, a closed form for block-diagonal B)."""
    g = load("rimp2_water_def2svp.npz")
    B, eo, ev = g["Bia"], g["eps_occ"], g["eps_vir"]
    e = O.rimp2(B, eo, ev)
    d = O.rimp2_dense(B, eo, ev)
    for k, f in (("EMP2_OS", "e_os"), ("EMP2_SS", "e_ss"), ("EMP2", "e_mp2"), ("EMP2_SCS", "e_scs")):
        assert abs(e[k] - float(g[f])) < 1e-14 and abs(d[k] - e[k]) < 1e-13
    assert -76.0 < float(g["e_rhf"]) < -75.9 and -0.21 < float(g["e_mp2"]) < -0.20
    assert int(g["naux_kept"]) == B.shape[0]
    from tests.test_gpu_posthf import _separated_waters
    Bn, eon, evn, ref = _separated_waters(3, seed=5)
    en = O.rimp2(Bn, eon, evn)
    for k in ref:
        assert abs(en[k] - ref[k]) < 1e-13
