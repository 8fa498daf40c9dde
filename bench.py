#!/usr/bin/env python3
"""RI-MP2 benchmark at the (H2O)32 / def2-SVP / def2-SVP-RI shape on one MI355X.

Workload (one "step"): B formation from the metric J and the 3-index
integrals (P|mu nu) already resident in HBM (Cholesky metric, overlapped on a
second stream with the AO->MO GEMMs, then the metric GEMM), then the fused
RI-MP2 pair-block energy, in one call (corr_ri_mp2_energy_from_j; on the
mixed path the metric GEMM writes B as the energy kernel's BF16 hi/lo
panels, so no FP64 B is stored or repacked).  Sizes are those of
(H2O)32: N = 768 basis functions, o = 160 occupied, v = 608 virtual,
naux = 2432 auxiliary functions.  Three paths are timed, K steps each:

  mixed: metric factorisation and the first AO->MO GEMM FP64 (FP64 MFMA),
         the second AO->MO GEMM and the Cholesky metric product as FP64
         GEMMs emulated on the int8 matrix cores (5 int8 slices per operand,
         exact int32 sums), energy BF16x3 MFMA with FP32 accumulation and FP64
         denominators/reduction
  fp64:  everything FP64 (FP64 MFMA)
  fp32:  B-formation GEMMs and energy in FP32 MFMA (reported beside, with
         its error against fp64)

FP32 has no place in the mixed path's B formation: the AO->MO sums cancel
(3e-6 rms relative error in (Q|ia) from FP32 GEMMs) and the Cholesky metric
(cond(J) = 1.1e6) amplifies that to ~1e-4 in B, an energy error of about
0.3 uEh per FP32 stage with a sign set by the summation order; an FP32
metric product alone costs 0.28 uEh.  The int8-emulated GEMMs keep B within
3e-7 of FP64 relative to its maximum (0.003 uEh).

value = algorithmic TFLOP/s of the mixed path, (W_B_chol + W_E) / t_step, with
W_E = 2 [o(o+1)/2] v^2 naux and W_B_chol = 2 naux o N^2 + 2 naux o v N + naux^2 o v
(the metric product with the triangular Cholesky factor; a dense J^-1/2 makes
it 2 naux^2 o v, W_B).

Inputs (--data real, the default) are real chemistry computed on the GPU
before the timed region: a seeded (H2O)32 cluster (rigid waters, O-H 0.958 A,
104.5 deg, random orientations, O atoms uniform in a cube of 40 A^3 per
molecule, O-O >= 2.8 A), its def2-SVP / def2-SVP-RI integrals, and an RI-JK
RHF (DIIS, |dE| <= 1e-10) whose orbitals and orbital energies feed the step.
--data synthetic uses seeded tensors of the same shapes instead.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-suite]

The line carries a "suite" object (--no-suite leaves it out) with the other rows of the hot path, each at its
own shape and on real chemistry: build_FMo and metric + B formation on the
(H2O)32 inputs above; AO unfold (of the unique list gathered from the exact
(mu nu|la si) tensor), 4-index transform, canonical and non-canonical MP2, the
Fock build and one RHF+DIIS iteration on the exact-integral RHF of (H2O)8
(N=192); the RI-MP2 gradient on the RI-JK RHF of (H2O)16; CCSD per iteration
and converged (FP64 and FP32) on the exact-integral RHF of (H2O)4.
--rows NAME,... restricts the suite to the named rows.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = "RI-MP2 s & TFLOP/s, (H2O)32/def2-SVP: mixed vs FP64 GPU at \u22641 \u00b5Eh error"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def sizes(nmol: int):
    N, o, naux = 24 * nmol, 5 * nmol, 76 * nmol
    return N, o, N - o, naux


def work(N, o, v, naux):
    """Algorithmic FLOPs: W_E of the pair-symmetric energy, and W_B of B
    formation (AO->MO products, then the metric product) with a dense metric
    J^-1/2.  The Cholesky path applies the triangular L^-1, whose product is
    half as much work: W_B_chol."""
    W_E = 2.0 * (o * (o + 1) // 2) * v * v * naux
    W_B = 2.0 * naux * o * N * N + 2.0 * naux * o * v * N + 2.0 * naux * naux * o * v
    W_B_chol = W_B - 1.0 * naux * naux * o * v
    return W_E, W_B, W_B_chol


def water_spectrum(nmol: int, rng: np.random.Generator):
    """Orbital energies shaped like def2-SVP water clusters: per molecule one
    core level near -20.5 Eh, four valence levels in [-1.4, -0.45] Eh and 19
    virtual levels in [0.15, 5.5] Eh, each with small seeded splittings."""
    core = -20.55 + 0.02 * rng.standard_normal(nmol)
    val = np.concatenate([np.array([-1.35, -0.72, -0.58, -0.50]) + 0.03 * rng.standard_normal(4)
                          for _ in range(nmol)])
    vir_base = np.array([0.18, 0.26, 0.80, 0.85, 1.05, 1.15, 1.25, 1.55, 1.70, 1.90, 2.25, 2.60,
                         3.00, 3.40, 3.90, 4.40, 4.80, 5.10, 5.50])
    vir = np.concatenate([vir_base + 0.04 * rng.standard_normal(19) for _ in range(nmol)])
    return np.sort(np.concatenate([core, val])), np.sort(vir)


def random_orthogonal(n: int, rng: np.random.Generator, device):
    """Seeded orthogonal n x n matrix D1 Q P Q D2 built from the orthonormal
    DCT-II matrix Q, a random permutation P and random sign diagonals D1, D2
    (closed form: no factorisation library call)."""
    import torch
    k = torch.arange(n, dtype=torch.float64, device=device)
    Q = torch.cos(torch.pi * (k[None, :] + 0.5) * k[:, None] / n) * (2.0 / n) ** 0.5
    Q[0] *= 0.5 ** 0.5
    perm = torch.from_numpy(rng.permutation(n)).to(device)
    d1 = torch.from_numpy(rng.choice([-1.0, 1.0], n)).to(device)
    d2 = torch.from_numpy(rng.choice([-1.0, 1.0], n)).to(device)
    return (d1[:, None] * (Q @ Q[perm]) * d2[None, :]).contiguous()


def make_inputs(nmol: int, seed: int, device):
    import torch
    N, o, v, naux = sizes(nmol)
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    rng = np.random.default_rng(seed)
    eo, ev = water_spectrum(nmol, rng)
    C = random_orthogonal(N, rng, device)
    Qa = random_orthogonal(naux, rng, device)
    lam = torch.logspace(-3, 3, naux, dtype=torch.float64, device=device)
    J = ((Qa * lam) @ Qa.T)
    J = (0.5 * (J + J.T)).contiguous()
    Pmn = torch.randn(naux, N, N, dtype=torch.float64, device=device, generator=g)
    Pmn = Pmn + Pmn.transpose(1, 2)
    return C, J, Pmn, torch.from_numpy(eo).to(device), torch.from_numpy(ev).to(device)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nmol", type=int, default=32, help="water molecules (shape only)")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="bounded CPU baseline budget")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 1 GiB device-to-device (a known-size kernel for scaling HBM counters)")
    ap.add_argument("--data", choices=("real", "synthetic"), default="real",
                    help="real: GPU integrals + RI-JK RHF of a seeded water cluster; synthetic: seeded tensors")
    ap.add_argument("--rows", default="", help="comma-separated suite rows to run (default: all)")
    ap.add_argument("--suite", dest="suite", action="store_true", default=True,
                    help="also time every other row of the hot path at its own shape (the 'suite' key; default)")
    ap.add_argument("--no-suite", dest="suite", action="store_false", help="headline step only")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported in this build "
              "(a multi-GPU run would shard the occupied pair list)", file=sys.stderr)
        return 2

    import torch
    import corr_amd
    from corr_amd._abi import Context

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.copy_calibration:
        x = torch.ones(1 << 27, dtype=torch.float64, device=dev)
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        del x, y
        log("calibration copy of 1 GiB done")
    N, o, v, naux = sizes(args.nmol)
    W_E, W_B, W_Bc = work(N, o, v, naux)
    log(f"(H2O){args.nmol}: N={N} o={o} v={v} naux={naux}; W_E={W_E/1e12:.3f} TF W_B={W_B/1e12:.3f} TF "
        f"(Cholesky metric: {W_Bc/1e12:.3f} TF)")
    inputs, real = {}, None
    if args.data == "real":
        from corr_amd import chem
        t_in = time.perf_counter()
        r = chem.water_cluster_rhf(args.nmol, args.seed, device=dev, log=log)
        C, J, Pmn, eo, ev = r["C"], r["J"], r["Pmn"], r["eps_occ"], r["eps_vir"]
        assert (r["nbf"], r["naux"], r["nocc"]) == (N, naux, o)
        scf = r["scf"]
        inputs = {"molecule": f"(H2O){args.nmol}, seed {args.seed}", "basis": "def2-SVP / def2-SVP-RI",
                  "rhf": "RI-JK (def2-SVP-RI), DIIS from iteration 8, |dE| <= 1e-10, start: superposition of each water's own RHF density",
                  "e_rhf": scf["e_total"], "rhf_iterations": scf["iterations"],
                  "rhf_converged": bool(scf["converged"]), "homo_lumo_Eh": float(ev[0] - eo[-1]),
                  "seconds_integrals": round(r["seconds_integrals"], 2), "seconds_rhf": round(r["seconds_scf"], 2),
                  "seconds_setup_total": round(time.perf_counter() - t_in, 2)}
        real = r
        del r
        torch.cuda.empty_cache()
        log(f"real inputs ready: E(RHF) = {scf['e_total']:.10f} ({scf['iterations']} iterations)")
    else:
        C, J, Pmn, eo, ev = make_inputs(args.nmol, args.seed, dev)
        # scale (P|mu nu) so that the FP64 correlation energy is ~ -0.2 Eh per water
        M, _ = corr_amd.ri_metric(J, "cholesky")
        B = corr_amd.ri_form_bia(C, Pmn, M, o, "cholesky", "fp64")
        e0 = corr_amd.rimp2_energy(B, eo, ev, "fp64")["EMP2"]
        Pmn.mul_((0.2 * args.nmol / abs(e0)) ** 0.25)
        del B, M
        torch.cuda.synchronize()
        log(f"synthetic inputs ready (unscaled E={e0:.4f})")
    ctx = Context.get(0)
    # The reference's metric (ri.cc:50-69) is J^-1/2 by eigendecomposition
    # with a cutoff; the step uses the Cholesky factor.  Both give the same
    # B B^T when no eigenvalue is cut: record the kept count and the FP64
    # energies of the two paths (outside the timed region).
    _, kept = corr_amd.ri_metric(J, "eigen")
    e_eig = corr_amd.rimp2_energy(corr_amd.ri_bia_from_j(C, Pmn, J, o, "eigen", "fp64")[0], eo, ev, "fp64")["EMP2"]
    e_chol = corr_amd.rimp2_energy(corr_amd.ri_bia_from_j(C, Pmn, J, o, "cholesky", "fp64")[0], eo, ev,
                                   "fp64")["EMP2"]
    torch.cuda.empty_cache()
    metric_info = {"step": "cholesky", "eigen_kept": kept, "naux": naux, "emp2_fp64_eigen": e_eig,
                   "emp2_fp64_cholesky": e_chol, "eigen_minus_cholesky_Eh": e_eig - e_chol}
    log(f"metric: eigen keeps {kept} of {naux}; FP64 EMP2 eigen - cholesky = {e_eig - e_chol:.3e} Eh")

    def step(prec: str):
        # metric factorisation overlapped with the AO->MO GEMMs, then the
        # energy; at bf16x3 the metric product writes the energy kernel's
        # BF16 hi/lo panels directly
        return corr_amd.ri_mp2_energy_from_j(C, Pmn, J, o, eo, ev, "cholesky", prec)[0]

    results = {}
    for name, prec in (("fp64", "fp64"), ("mixed", "bf16x3"), ("fp32", "fp32")):
        for _ in range(args.warmup):
            step(prec)
        torch.cuda.synchronize()
        ctx.kernel_time(None)
        ctx.enable_timing(True)
        t0 = time.perf_counter()
        for k in range(args.steps):
            e = step(prec)
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / args.steps
        ms, n = ctx.kernel_time("rimp2_energy")
        ctx.enable_timing(False)
        results[name] = {"s_per_step": t, "energy": e, "kernel_ms": ms / max(n, 1), "launches": n}
        log(f"{name}: {t*1e3:.2f} ms/step, energy kernel {ms/max(n,1):.2f} ms, EMP2={e['EMP2']:.15f}")

    mixed, f64 = results["mixed"], results["fp64"]
    err_uEh = abs(mixed["energy"]["EMP2"] - f64["energy"]["EMP2"]) * 1e6
    t_k = mixed["kernel_ms"] / 1e3
    achieved = 3.0 * W_E / t_k / 1e12
    t_k64 = f64["kernel_ms"] / 1e3

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(C, Pmn, J, o, eo, ev, W_E + W_Bc, args.cpu_seconds)

    line = {
        "metric": METRIC,
        "value": round((W_E + W_Bc) / mixed["s_per_step"] / 1e12, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed["s_per_step"] * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3 energy (fp32 accumulation; B formation fp64 and int8-emulated fp64 GEMMs, "
                 "fp64 denominators and reduction)",
        "data": ("real chemistry: seeded water cluster, def2-SVP/def2-SVP-RI integrals and RI-JK RHF computed on the "
                 "GPU by this library before the timed region" if args.data == "real" else
                 "synthetic (seeded tensors of the (H2O)32 shapes; no integral files)"),
        "inputs": inputs or None,
        "config": {"workload": f"RI-MP2 B formation + energy, (H2O){args.nmol} def2-SVP/def2-SVP-RI shape",
                   "N": N, "nocc": o, "nvir": v, "naux": naux, "parallelism": "single GPU"},
        "mixed_s": round(mixed["s_per_step"], 5),
        "fp64_s": round(f64["s_per_step"], 5),
        "speedup_mixed_vs_fp64": round(f64["s_per_step"] / mixed["s_per_step"], 3),
        "error_uEh": round(err_uEh, 4),
        "emp2_fp64": f64["energy"]["EMP2"],
        "emp2_mixed": mixed["energy"]["EMP2"],
        "fp64_tflops": round((W_E + W_Bc) / f64["s_per_step"] / 1e12, 3),
        "work_TFLOP": {"W_E": round(W_E / 1e12, 4), "W_B_cholesky": round(W_Bc / 1e12, 4),
                       "W_B_dense_metric": round(W_B / 1e12, 4),
                       "note": "value = (W_E + W_B_cholesky) / t; the step applies the triangular Cholesky factor"},
        "ri_metric": metric_info,
        "fp32": {"s": round(results["fp32"]["s_per_step"], 5),
                 "speedup_vs_fp64": round(f64["s_per_step"] / results["fp32"]["s_per_step"], 3),
                 "error_uEh": round(abs(results["fp32"]["energy"]["EMP2"] - f64["energy"]["EMP2"]) * 1e6, 4),
                 "note": "B formation and energy in FP32 MFMA (FP32 accumulation, FP64 denominators/reduction)"},
        "roofline": {"bound": "mfma", "kernel": "rimp2_energy_bf16x3_wide", "achieved": round(achieved, 2),
                     "peak": 2500.0, "unit": "TFLOP/s", "frac": round(achieved / 2500.0, 4),
                     "traffic": recorded_traffic("rimp2_energy_bf16x3_wide"),
                     "traffic_source": "profiles/rimp2_energy_h2o32.md (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE; "
                                       "scaled by a 1 GiB copy)",
                     "work": "3 x W_E (three BF16 products per FP32 product)"},
        "roofline_fp64": {"bound": "mfma", "kernel": "rimp2_energy_real<double>",
                          "achieved": round(W_E / t_k64 / 1e12, 2), "peak": 78.6, "unit": "TFLOP/s",
                          "frac": round(W_E / t_k64 / 1e12 / 78.6, 4)},
        "cpu_baseline": cpu,
    }
    if args.suite:
        keep = {"C": C, "J": J, "Pmn": Pmn, "F": real["F"] if real is not None else None, "nocc": o,
                "S": real["S"] if real is not None else None, "mol": real["mol"] if real is not None else None}
        del C, J, Pmn, real
        torch.cuda.empty_cache()
        line["suite"] = run_suite(dev, args.seed, [r for r in args.rows.split(",") if r], keep)
    print(json.dumps(line), flush=True)
    return 0


# ---------------------------------------------------------------------------
# Suite: the other rows of the hot path, each at its own shape.
# ---------------------------------------------------------------------------

def _timed(fn, reps: int = 3):
    """Median wall time (s) of fn() over reps runs after one warm-up, with the
    device synchronised around each run."""
    import torch
    fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
        del out
    return float(np.median(ts))


def run_suite(dev, seed: int, rows=(), main=None) -> dict:
    """Every other row at its own shape, on real chemistry computed on the GPU
    (seeded water clusters, def2-SVP / def2-SVP-RI, exact-integral RHF up to
    (H2O)8, RI-JK RHF above).  `main` holds the headline (H2O)32 inputs."""
    import torch
    import corr_amd
    from corr_amd import chem
    from corr_amd._abi import Context
    from corr_amd.pipeline import sorted_unique_list
    out = {"peaks_TFLOPs": {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}}

    def rec(name, fn):
        if rows and name not in rows:
            return
        try:
            out[name] = fn()
            log(f"suite {name}: {out[name]}")
        except Exception as exc:   # a failing row is reported, not fatal
            out[name] = {"error": f"{type(exc).__name__}: {exc}"}
            log(f"suite {name}: FAILED {exc}")
        torch.cuda.empty_cache()

    rng = np.random.default_rng(seed)
    ctx = Context.get(dev.index if dev.index is not None else torch.cuda.current_device())

    # a1: FMo = C F C^T at N = 768 (the headline inputs)
    def fmo():
        if main is not None and main.get("F") is not None:
            Cm, F, src = main["C"], main["F"], "real (H2O)32 RHF"
        else:
            Cm, F, src = random_orthogonal(768, rng, dev), torch.randn(768, 768, dtype=torch.float64, device=dev), \
                "synthetic"
        N = Cm.shape[0]
        t = _timed(lambda: corr_amd.build_FMo(Cm, F))
        return {"shape": f"N={N}", "inputs": src, "ms": round(t * 1e3, 3), "TFLOPs": round(4 * N ** 3 / t / 1e12, 2)}
    rec("build_fmo_h2o32", fmo)

    # orbital localisation (pymodule.py:75-167) of the 160 occupied RHF orbitals of (H2O)32
    def localize():
        if main is None or main.get("S") is None:
            return {"skipped": "needs the real (H2O)32 inputs"}
        from corr_amd import integrals as GI
        from corr_amd.basis import ShellArrays, read_g94
        Cocc = main["C"][:main["nocc"]].contiguous()
        O = ShellArrays(main["mol"], read_g94("def2-svp"))
        D = GI.dipole(O)
        res = {"shape": f"(H2O)32 nocc={Cocc.shape[0]} nbf={Cocc.shape[1]} atoms={len(main['mol'].symbols)}"}
        for kind in ("pm", "boys"):
            t0 = time.perf_counter()
            info, _ = corr_amd.localize(Cocc, kind, S=main["S"], bf_atom=O.bf_atom, D=D)
            torch.cuda.synchronize()
            res[f"{kind}_s"] = round(time.perf_counter() - t0, 3)
            res[f"{kind}_sweeps"] = info["sweeps"]
            res[f"{kind}_converged"] = info["converged"]
            res[f"{kind}_spread"] = [round(info["spread_initial"], 6), round(info["spread"], 6)]
        return res
    rec("localize_h2o32", localize)

    # (H2O)8: exact integrals and RHF, then unfold + 4-index + canonical and non-canonical MP2 + Fock build
    h2o8 = {}

    def fourindex():
        r = chem.water_cluster_rhf(8, seed, jk="dense", log=log)
        h2o8.update(r)
        N, o = r["nbf"], r["nocc"]
        G = r["G"]
        val, nums, sc = sorted_unique_list(G)
        n_unique = val.numel()
        t_unf = _timed(lambda: corr_amd.ao_unfold(N, val, nums, sc), reps=2)
        dense = corr_amd.ao_unfold(N, val, nums, sc)
        # the engine computes (ab|cd) and (cd|ab) separately: equal to rounding, not bitwise
        unf_diff = float((dense - G).abs().max().item())
        nbytes_unf = n_unique * 16 + dense.numel() * 8
        del val, nums, dense
        torch.cuda.empty_cache()
        Cm = r["C"]
        FMo = corr_amd.build_FMo(Cm, r["F"])
        # four full quarter transforms: 8 N^5; executed: 6 N^5 (quarters 3 and 4
        # run on the packed pairs i <= j of the (i, j)-symmetric half transform)
        W4, W4x = 8.0 * N ** 5, 6.0 * N ** 5
        res = {"shape": "(H2O)8 N=192 o=40, real integrals", "e_rhf": r["scf"]["e_total"],
               "rhf_iterations": r["scf"]["iterations"], "unique_integrals": n_unique,
               "unfold_ms": round(t_unf * 1e3, 2), "unfold_GBs": round(nbytes_unf / t_unf / 1e9, 1),
               "unfold_vs_tensor_max_abs": unf_diff, "W_4_full_quarters_TFLOP": round(W4 / 1e12, 3),
               "W_4_executed_TFLOP": round(W4x / 1e12, 3)}
        emp2 = {}
        for prec in ("fp64", "fp32"):
            t = _timed(lambda: corr_amd.MOtrans(Cm, G, prec), reps=2)
            res[f"motrans_{prec}_ms"] = round(t * 1e3, 2)
            res[f"motrans_{prec}_executed_TFLOPs"] = round(W4x / t / 1e12, 2)
            ints = corr_amd.MOtrans(Cm, G, prec)
            emp2[prec] = corr_amd.run_canonical_mp2(ints, FMo, o)["EMP2"]
            if prec == "fp64":
                t = _timed(lambda: corr_amd.run_canonical_mp2(ints, FMo, o))
                res["canonical_mp2_ms"] = round(t * 1e3, 3)
                res["canonical_mp2_GBs"] = round(2 * o * o * (N - o) ** 2 * 8 / t / 1e9, 1)
                t1 = _timed(lambda: corr_amd.run_non_canonical_mp2(ints, FMo, o, 1), reps=2)
                t21 = _timed(lambda: corr_amd.run_non_canonical_mp2(ints, FMo, o, 21), reps=2)
                it = (t21 - t1) / 20
                ov2 = o * o * (N - o) ** 2
                res["non_canonical_mp2_iteration_ms"] = round(it * 1e3, 3)
                # per iteration: G GEMM reads T, writes G; the step reads K, T, G and both transposes, writes T
                res["non_canonical_mp2_GBs"] = round(8 * ov2 * 8 / it / 1e9, 1)
            del ints
            torch.cuda.empty_cache()
        res["emp2_fp64"] = emp2["fp64"]
        res["emp2_fp32_error_uEh"] = round(abs(emp2["fp32"] - emp2["fp64"]) * 1e6, 4)
        return res
    rec("fourindex_h2o8", fourindex)

    # RHF input side at (H2O)8: Fock build from the dense tensor (HBM-bound:
    # slabs p >= q read once) and one RHF + DIIS iteration
    def rhf():
        r = h2o8 if h2o8 else chem.water_cluster_rhf(8, seed, jk="dense", log=log)
        N, nroe = r["nbf"], r["nroe"]
        G, H, S = r["G"], r["H"], r["S"]
        P = corr_amd.build_Pmat(r["C"], nroe)
        ctx.enable_timing(True)
        ctx.kernel_time(None)
        t = _timed(lambda: corr_amd.build_Fmat(G, P, H), reps=5)
        ms, n = ctx.kernel_time("fock_slab_kernel")
        ctx.enable_timing(False)
        slab_bytes = N * (N + 1) // 2 * N * N * 8
        res = {"shape": "(H2O)8 N=192 nroe=80, real integrals", "fock_ms": round(t * 1e3, 3),
               "fock_slab_kernel_ms": round(ms / max(n, 1), 3),
               "fock_slab_bytes_GB": round(slab_bytes / 1e9, 3),
               "fock_slab_GBs": round(slab_bytes / (ms / max(n, 1) * 1e-3) / 1e9, 1)}
        C0 = r["C"]
        t1 = _timed(lambda: corr_amd.run_scf(H, S, G, C0, nroe, maxiter=1, e_tol=0.0, core_guess_check=False),
                    reps=2)
        t9 = _timed(lambda: corr_amd.run_scf(H, S, G, C0, nroe, maxiter=9, e_tol=0.0, core_guess_check=False),
                    reps=2)
        res["rhf_iteration_ms"] = round((t9 - t1) / 8 * 1e3, 2)
        res["rhf_setup_ms"] = round((t1 - (t9 - t1) / 8) * 1e3, 2)
        return res
    rec("rhf_h2o8", rhf)
    h2o8.clear()
    torch.cuda.empty_cache()

    # a7-a8: metric and B formation at (H2O)32 (the headline inputs)
    def bform():
        if main is not None:
            C, J, Pmn = main["C"], main["J"], main["Pmn"]
            src = "real (H2O)32"
        else:
            C, J, Pmn, _, _ = make_inputs(32, seed, dev)
            src = "synthetic"
        N, o, v, naux = sizes(32)
        res = {"shape": "(H2O)32 N=768 naux=2432", "inputs": src}
        res["metric_eigen_ms"] = round(_timed(lambda: corr_amd.ri_metric(J, "eigen")) * 1e3, 2)
        res["metric_cholesky_ms"] = round(_timed(lambda: corr_amd.ri_metric(J, "cholesky")) * 1e3, 2)
        M, _ = corr_amd.ri_metric(J, "cholesky")
        _, WB, WBc = work(N, o, v, naux)
        res["W_B_cholesky_TFLOP"] = round(WBc / 1e12, 4)
        res["W_B_dense_metric_TFLOP"] = round(WB / 1e12, 4)
        # bf16x3: the mixed path's B formation (FP64 AO->MO, int8-emulated FP64 metric product)
        for prec in ("fp64", "bf16x3", "fp32"):
            t = _timed(lambda: corr_amd.ri_form_bia(C, Pmn, M, o, "cholesky", prec))
            res[f"form_bia_{prec}_ms"] = round(t * 1e3, 2)
            # the triangular metric product is credited at its own (half) work
            res[f"form_bia_{prec}_TFLOPs"] = round(WBc / t / 1e12, 2)
            t = _timed(lambda: corr_amd.ri_bia_from_j(C, Pmn, J, o, "cholesky", prec))
            res[f"bia_from_J_overlapped_{prec}_ms"] = round(t * 1e3, 2)
        return res
    rec("b_formation_h2o32", bform)

    # RI-MP2 gradient at (H2O)16 on its RI-JK RHF
    def grad():
        r = chem.water_cluster_rhf(16, seed, log=log)
        N, o, v, naux = r["nbf"], r["nocc"], r["nbf"] - r["nocc"], r["naux"]
        B, _ = corr_amd.ri_bia_from_j(r["C"], r["Pmn"], r["J"], o, "cholesky", "fp64")
        eo, ev = r["eps_occ"], r["eps_vir"]
        del r
        torch.cuda.empty_cache()
        WE = work(N, o, v, naux)[0]
        res = {"shape": "(H2O)16 o=80 v=304 naux=1216, real B", "W_E_TFLOP": round(WE / 1e12, 4)}
        grads = {}
        for prec in ("fp64", "fp32"):
            te = _timed(lambda: corr_amd.rimp2_energy(B, eo, ev, prec))
            tg = _timed(lambda: corr_amd.rimp2_gradient(B, eo, ev, 1.0, 1.0, prec))
            res[f"energy_{prec}_ms"] = round(te * 1e3, 2)
            res[f"gradient_{prec}_ms"] = round(tg * 1e3, 2)
            res[f"gradient_{prec}_TFLOPs"] = round(3 * WE / tg / 1e12, 2)
            grads[prec] = corr_amd.rimp2_gradient(B, eo, ev, 1.0, 1.0, prec)
        E64, dB64 = grads["fp64"][0], grads["fp64"][1]
        E32, dB32 = grads["fp32"][0], grads["fp32"][1]
        res["emp2_fp64"] = E64
        res["fp32_energy_error_uEh"] = round(abs(E32 - E64) * 1e6, 4)
        res["fp32_dE_dB_max_rel_error"] = float(((dB32 - dB64).abs().max() / dB64.abs().max()).item())
        return res
    rec("rimp2_gradient_h2o16", grad)

    # a10-a19: CCSD at (H2O)4 on its exact-integral RHF (spin orbitals: o = 40, v = 152)
    def ccsd():
        r = chem.water_cluster_rhf(4, seed, jk="dense", log=log)
        N, nroe = r["nbf"], r["nroe"]
        o, v = nroe, 2 * N - nroe
        prec_ints = corr_amd.MOtrans(r["C"], r["G"], "fp64")
        Cm, h, ion = r["C"], r["H"], r["mol"].nuclear_repulsion()
        # work of this implementation: the pp ladder over packed antisymmetric pairs
        # 2 Po Pv^2 (Po = o(o-1)/2, Pv = v(v-1)/2), ring 2 x 2 (ov)^3, Wmnij tau term and hh
        # ladder over packed pairs 2 x 2 Po^2 Pv; the reference's intermediates:
        # 4 o^2 v^4 + 4 o^3 v^3 + 4 o^4 v^2 (unpacked ladders, and Wabef's tau term, carried
        # here by Wmnij)
        Po, Pv = o * (o - 1) // 2, v * (v - 1) // 2
        W = 2.0 * Po * Pv * Pv + 4.0 * o ** 3 * v ** 3 + 4.0 * Po * Po * Pv
        W_ref = 4.0 * o * o * v ** 4 + 4.0 * o ** 3 * v ** 3 + 4.0 * o ** 4 * v * v
        # executed with spin-symmetric amplitudes: each ring product as one 2x2-class and two
        # cross-class GEMMs (20 B4^3 each, B4 = o v / 4), the pp ladder, Wmnij's tau term and
        # the hh ladder on their three pair spin classes (alpha alpha, beta beta: h(h-1)/2
        # pairs; mixed: h^2 pairs)
        oh, vh, B4 = o // 2, v // 2, o * v // 4
        po_c, pv_c = [oh * (oh - 1) // 2] * 2 + [oh * oh], [vh * (vh - 1) // 2] * 2 + [vh * vh]
        W_exec = (2 * 20.0 * B4 ** 3 + sum(2.0 * a * b * b for a, b in zip(po_c, pv_c))
                  + sum(4.0 * a * a * b for a, b in zip(po_c, pv_c)))
        res = {"shape": "(H2O)4 N=96 spin-orbital o=40 v=152, real integrals",
               "e_rhf": r["scf"]["e_total"], "W_CC_executed_TFLOP_per_iter": round(W_exec / 1e12, 3),
               "W_CC_dense_spin_orbital_TFLOP": round(W / 1e12, 3),
               "W_CC_reference_intermediates_TFLOP": round(W_ref / 1e12, 3)}
        del r
        energies = {}
        for prec in ("fp64", "fp32"):
            t1 = _timed(lambda: corr_amd.ccsd_ur(prec_ints, Cm, h, nroe, ion, maxiter=1, prec=prec), reps=1)
            t11 = _timed(lambda: corr_amd.ccsd_ur(prec_ints, Cm, h, nroe, ion, maxiter=11, prec=prec), reps=1)
            it = (t11 - t1) / 10
            res[f"iteration_{prec}_ms"] = round(it * 1e3, 2)
            res[f"iteration_{prec}_executed_TFLOPs"] = round(W_exec / it / 1e12, 2)
            res[f"setup_{prec}_ms"] = round((t1 - it) * 1e3, 1)
            c = corr_amd.ccsd_ur(prec_ints, Cm, h, nroe, ion, maxiter=100, mode="diis", prec=prec, e_tol=1e-9,
                                 r_tol=1e-7)
            energies[prec] = c["e_ccsd"]
            res[f"converged_{prec}_iterations"] = c["iterations"]
        res["e_ccsd_fp64"] = energies["fp64"]
        res["fp32_error_uEh"] = round(abs(energies["fp32"] - energies["fp64"]) * 1e6, 4)
        return res
    rec("ccsd_h2o4", ccsd)
    return out


def recorded_traffic(kernel: str):
    """HBM bytes per launch of `kernel` as recorded from the FETCH_SIZE and
    WRITE_SIZE counter passes in profiles/rimp2_energy_h2o32.md (a line
    'traffic_bytes_per_launch <kernel>: <bytes>'); None if not recorded."""
    import re
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "rimp2_energy_h2o32.md")
    try:
        with open(path) as fh:
            m = re.search(r"traffic_bytes_per_launch " + re.escape(kernel) + r":\s*(\d+)", fh.read())
    except OSError:
        return None
    return int(m.group(1)) if m else None


def cpu_baseline(C, Pmn, J, o, eo, ev, W, budget_s):
    """Times the oracle's restatement of the step's region in the reference's
    own order on the host, each part on a bounded sample, extrapolated:
      J^-1/2 by eigendecomposition with the cutoff (ri.cc:50-69), whole;
      BPQ = J^-1/2 (P|mu nu) (ri.cc:102-103), one dgemm, on a slice of mu;
      transform_ri (ri.cc:147-172) on a sample of Q slices;
      run_canonical_mp2_ri (can_mp2.cc:82-112), one dgemm per ordered pair,
      on a sample of pairs (about 60% of the budget).
    numpy / OpenBLAS on all host threads (the reference runs transform_ri on
    one thread).  value = W / t with the GPU step's algorithmic W."""
    import torch
    import corr_amd
    from oracle import posthf as O
    try:
        from threadpoolctl import threadpool_info
        threads = max((p.get("num_threads", 1) for p in threadpool_info()
                       if p.get("internal_api") == "openblas"), default=os.cpu_count())
    except Exception:
        threads = os.cpu_count()
    naux, N = Pmn.shape[0], Pmn.shape[1]
    v = N - o
    parts = {}
    Jh = J.cpu().numpy()
    t0 = time.perf_counter()
    Mh, _ = O.ri_metric_eigen(Jh)
    parts["metric_eigen"] = (time.perf_counter() - t0, 1.0, "whole")
    # BPQ on a slice of mu rows
    s_mu = 24
    P_s = Pmn[:, :s_mu, :].contiguous().cpu().numpy()
    t0 = time.perf_counter()
    Mh @ P_s.reshape(naux, -1)   # the dgemm of O.ri_bpq (ri.cc:102-103) on s_mu rows of (P|mu nu)
    parts["bpq_dgemm"] = (time.perf_counter() - t0, N / s_mu, f"{s_mu} of {N} mu rows")
    del P_s
    # transform_ri on Q slices of the real BPQ
    nq = 96
    Mg = torch.from_numpy(Mh[:nq]).to(Pmn.device)
    BPQ_s = (Mg @ Pmn.view(naux, -1)).view(nq, N, N).cpu().numpy()
    Ch = C.cpu().numpy()
    t0 = time.perf_counter()
    O.transform_ri(Ch, BPQ_s, o)
    parts["transform_ri"] = (time.perf_counter() - t0, naux / nq, f"{nq} of {naux} Q slices")
    del BPQ_s, Mg
    # energy on a sample of ordered pairs of the real B
    M, _ = corr_amd.ri_metric(J, "cholesky")
    Bia = corr_amd.ri_form_bia(C, Pmn, M, o, "cholesky", "fp64").cpu().numpy()
    eo_h, ev_h = eo.cpu().numpy(), ev.cpu().numpy()
    rng = np.random.default_rng(0)
    pairs_all = [(i, j) for i in range(o) for j in range(o)]
    order = rng.permutation(len(pairs_all))
    O.rimp2(Bia, eo_h, ev_h, pairs=[pairs_all[order[0]]])   # warm-up
    budget_e = max(budget_s - sum(p[0] for p in parts.values()), 0.3 * budget_s)
    done, batch = 0, 16
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_e and done < len(pairs_all):
        sel = [pairs_all[k] for k in order[done:done + batch]]
        O.rimp2(Bia, eo_h, ev_h, pairs=sel)
        done += len(sel)
    parts["rimp2_energy"] = (time.perf_counter() - t0, len(pairs_all) / done, f"{done} of {len(pairs_all)} pairs")
    sampled = sum(p[0] for p in parts.values())
    full = {k: p[0] * p[1] for k, p in parts.items()}
    total = sum(full.values())
    log("cpu baseline: " + ", ".join(f"{k} {parts[k][0]:.1f} s ({parts[k][2]}) -> {full[k]:.1f} s" for k in parts)
        + f"; step {total:.1f} s")
    return {"value": round(W / total / 1e12, 5), "unit": "TFLOP/s", "cores": int(threads), "kind": "port",
            "sample": "oracle restatement of the step in the reference's order (eigen metric ri.cc:50-69; "
                      "BPQ dgemm ri.cc:102-103; transform_ri ri.cc:147-172; per-pair dgemm energy "
                      "can_mp2.cc:82-112), numpy/OpenBLAS on the host: "
                      + "; ".join(f"{k}: {parts[k][2]} in {parts[k][0]:.2f} s" for k in parts)
                      + f"; {sampled:.1f} s sampled, step extrapolated to {total:.0f} s",
            "seconds_full_step_extrapolated": round(total, 1),
            "seconds_by_part_extrapolated": {k: round(x, 2) for k, x in full.items()}}


if __name__ == "__main__":
    sys.exit(main())
