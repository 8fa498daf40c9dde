"""File-driven run of the reference's program (src/main.cc:27-240) on the GPU.

    read_system -> calc_ion_rep -> read_oei -> read_tei -> calc_S12 ->
    read_wav_HF -> check_initial_guess -> run_scf -> build_FMo -> MOtrans ->
    run_canonical_mp2 -> (BPQ) -> transform_ri -> run_canonical_mp2_ri -> ccsd_ur

Inputs are the psi4 plugin's files PREFIX.{sys,oei,tei,ahfw[,rimp2]} (layouts
in corr_amd/formats.py).  Two steps of main.cc need libint2 and have no file
form here: the two-electron list when .sys says nrofint == 0
(calculate_libint_tei, main.cc:96-103) and the RI integrals (calculate_ri,
main.cc:146).  For the first the run stops with an error; for the second
PREFIX.rimp2 (the BPQ the plugin exports, read as read_transform_ri does,
ri.cc:109-145) takes its place, and the RI-MP2 step is skipped without it.
"""
from __future__ import annotations

import os
import time

import numpy as np
import torch

from . import formats as FM
from .ccsd import ccsd_ur
from .posthf import MOtrans, ao_unfold, build_FMo, run_canonical_mp2, run_canonical_mp2_ri, transform_ri
from .scf import run_scf


def _gpu(x, device):
    return torch.from_numpy(np.array(x, dtype=np.float64, copy=True)).to(device)


def sorted_unique_list(dense: torch.Tensor):
    """The reference's sorted unique integral list (TEints: intval, intnums,
    sortcount; util.h:61-65) of a dense (mu nu|la si) tensor on the GPU.

    Every canonical quadruple (a >= b, c >= d, ab >= cd) is put in the
    reference's stored form by resort step 1 (ops_rhf.cc:766-818), classified
    into its five permutation classes (ops_rhf.cc:820-896: all four equal;
    (aa|cc); (ab|ab); (aa|cd); the rest) and stably sorted by class.  Values
    are gathered from `dense`; nothing is screened.  Returns (intval float64
    [n], intnums int16 [n, 4], sortcount [4] as Python ints)."""
    N = dense.shape[0]
    device = dense.device
    P = N * (N + 1) // 2
    T = P * (P + 1) // 2
    k = torch.arange(T, dtype=torch.int64, device=device)

    def tri(x):
        r = torch.floor((torch.sqrt(8.0 * x.double() + 1.0) - 1.0) / 2.0).long()
        r = torch.where(r * (r + 1) // 2 > x, r - 1, r)
        r = torch.where((r + 1) * (r + 2) // 2 <= x, r + 1, r)
        return r, x - r * (r + 1) // 2

    p, q = tri(k)
    del k
    a, b = tri(p)
    c, d = tri(q)
    del p, q
    oa, ob, oc, od = a.clone(), b.clone(), c.clone(), d.clone()

    def assign(m, w, x, y, z):
        a[m], b[m], c[m], d[m] = w[m], x[m], y[m], z[m]

    assign((oa == ob) & (ob == od) & (oc != od), oa, oa, oa, oc)
    assign((oa != ob) & (oa == oc) & (oc == od), oa, oa, oa, ob)
    assign((oa != ob) & (ob == oc) & (oc == od), ob, ob, ob, oa)
    assign((oa == od) & (ob == oc) & (oa != ob), oa, ob, od, oc)
    assign((ob == oc) & (ob != oa) & (ob != od) & (oa != od), ob, oa, oc, od)
    assign((oc == od) & (oc != oa) & (oc != ob) & (oa != ob), oc, oc, oa, ob)
    assign((oa == od) & (oa != ob) & (oa != oc) & (ob != oc), oa, ob, od, oc)
    assign((ob == od) & (ob != oa) & (ob != oc) & (oa != oc), ob, oa, od, oc)
    del oa, ob, oc, od
    cls = torch.full_like(a, 4)
    m0 = (a == b) & (b == c) & (c == d)
    m1 = ~m0 & (a == b) & (c == d)
    m2 = ~m0 & ~m1 & (a == c) & (b == d)
    m3 = ~m0 & ~m1 & ~m2 & (a == b)
    cls[m3] = 3
    cls[m2] = 2
    cls[m1] = 1
    cls[m0] = 0
    order = torch.sort(cls, stable=True).indices
    sortcount = [int((cls <= j).sum()) for j in range(4)]
    nums = torch.stack([a, b, c, d], dim=1)[order].to(torch.int16)
    del a, b, c, d, cls, order
    q = nums.to(torch.int64)
    val = dense[q[:, 0], q[:, 1], q[:, 2], q[:, 3]].contiguous()
    del q
    return val, nums.contiguous(), sortcount


def run_from_files(prefix: str, prec: str = "fp64", ri_prec: str | None = None, scf_maxiter: int = 100,
                   ccsd_maxiter: int = 20, ccsd: bool = True, device: str | int | None = None,
                   log=print) -> dict:
    """Runs main.cc's sequence on PREFIX.* and returns every energy it prints.

    prec is the arithmetic of the 4-index transform, ri_prec (default prec)
    of transform_ri and the RI-MP2 energy.  log (None to silence) receives
    the reference's result lines."""
    say = log or (lambda *a, **k: None)
    t0 = time.perf_counter()
    sysinfo = FM.read_system(prefix + ".sys")
    N, nroe = sysinfo.nroao, sysinfo.nroe
    if nroe <= 0 or nroe % 2 or nroe // 2 > N:
        raise ValueError(f"{prefix}.sys: closed-shell RHF needs an even electron count <= 2 nroao (nroe={nroe})")
    ion_rep = sysinfo.ion_rep
    oei = FM.read_oei(prefix + ".oei", N)
    if sysinfo.nrofint <= 0:
        raise ValueError(f"{prefix}.sys: nrofint = 0 asks for libint2 integrals, which this build does not "
                         "evaluate; export the .tei file with the plugin")
    sortcount, intval, intnums = FM.read_tei(prefix + ".tei", sysinfo.nrofint)
    _, MOs0 = FM.read_wav(prefix + ".ahfw", N)
    out = {"nroao": N, "nroe": nroe, "ion_rep": ion_rep, "basis": sysinfo.basis_names}
    if device is None:
        device = torch.cuda.current_device()
    dev = torch.device("cuda", device) if isinstance(device, int) else torch.device(device)

    H, S = _gpu(oei["H"], dev), _gpu(oei["S"], dev)
    idx = torch.from_numpy(np.array(intnums, dtype=np.uint16).view(np.int16)).to(dev)
    dense = ao_unfold(N, _gpu(intval, dev), idx, sortcount)
    del idx
    scf = run_scf(H, S, dense, _gpu(MOs0, dev), nroe, ion_rep=ion_rep, maxiter=scf_maxiter)
    out["scf"] = {k: scf[k] for k in ("e_scf", "e_one", "e_two", "e_total", "iterations", "converged", "diis",
                                      "core_guess")}
    out["scf"]["energies"] = scf["energies"]
    say(f"{'  ONE E:':>12}{scf['e_one']:16.10f}")
    say(f"{'  TWO E:':>12}{scf['e_two']:16.10f}")
    say(f"{'TOTAL E:':>12}{scf['e_total']:16.10f}")
    C, F = scf["MOs"], scf["Fmat"]
    out["MOens"] = scf["MOens"].cpu().numpy()

    FMo = build_FMo(C, F)
    nocc = nroe // 2
    prec_ints = MOtrans(C, dense, prec)
    del dense
    mp2 = run_canonical_mp2(prec_ints, FMo, nocc)
    out["mp2"] = mp2
    say(f"{'EMP2:':>12}{mp2['EMP2']:16.10f}")

    if os.path.exists(prefix + ".rimp2"):
        rp = ri_prec or prec
        BPQ = torch.from_numpy(np.array(FM.read_bpq(prefix + ".rimp2", sysinfo.naux_2, N))).to(dev)
        Bia = transform_ri(C, BPQ, nocc, rp)
        del BPQ
        rimp2 = run_canonical_mp2_ri(Bia, FMo, nocc, rp)
        out["rimp2"] = rimp2
        say(f"{'EMP2(RI):':>12}{rimp2['EMP2']:16.10f}")

    if ccsd:
        cc = ccsd_ur(prec_ints, C, H, nroe, ion_rep, maxiter=ccsd_maxiter)
        out["ccsd"] = {k: cc[k] for k in ("energies", "emp2", "e_hf_check", "e_ccsd", "iterations")}
        say(f"{'ECCSD:':>12}{cc['e_ccsd']:16.10f}")
    out["seconds"] = time.perf_counter() - t0
    return out
