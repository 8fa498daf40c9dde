"""Runs the GEMM-bound rows once at their suite shapes on synthetic inputs,
for counter passes (rocprofv3 --pmc ...): a 1 GiB device copy (calibration
of FETCH_SIZE / WRITE_SIZE), the 4-index transform of a dense N = 192 tensor
and canonical MP2 (o = 40), two CCSD updates at N = 96 with 40 electrons,
and FP32 B formation at the (H2O)32 shape (N = 768, o = 160, naux = 2432).
The values do not change the kernels' work.

    python3 -m corr_amd.profile_rows
"""
import torch

import corr_amd


def main() -> None:
    print("profile rows", flush=True)
    dev = torch.device("cuda", 0)
    xc = torch.ones(1 << 27, dtype=torch.float64, device=dev)
    yc = torch.empty_like(xc)
    yc.copy_(xc)
    torch.cuda.synchronize()
    del xc, yc
    print("calibration copy done", flush=True)
    g = torch.Generator(device=dev)
    g.manual_seed(3)
    # 4-index transform + canonical MP2 at N = 192, o = 40
    N, o = 192, 40
    C, _ = torch.linalg.qr(torch.randn(N, N, dtype=torch.float64, device=dev, generator=g))
    C = C.contiguous()
    X = torch.randn(3 * N, N, N, dtype=torch.float64, device=dev, generator=g) / N
    X = 0.5 * (X + X.transpose(1, 2))
    G = torch.einsum("Pab,Pcd->abcd", X, X).contiguous()
    del X
    ints = corr_amd.MOtrans(C, G, "fp64")
    FMo = torch.diag(torch.cat([torch.linspace(-2, -0.5, o), torch.linspace(0.2, 3, N - o)]).to(dev, torch.float64)).contiguous()
    e = corr_amd.run_canonical_mp2(ints, FMo, o)
    print("motrans + mp2", e["EMP2"], flush=True)
    del G, ints
    torch.cuda.empty_cache()
    # CCSD at N = 96 (nroe 40): 2 Jacobi updates
    N, nroe = 96, 40
    h = torch.diag(torch.linspace(-2.0, 1.5, N, dtype=torch.float64, device=dev))
    X = torch.randn(2 * N, N, N, dtype=torch.float64, device=dev, generator=g) / N
    X = 0.5 * (X + X.transpose(1, 2))
    eri = (0.05 * torch.einsum("Pab,Pcd->abcd", X, X)).contiguous()
    del X
    r = corr_amd.ccsd_ur(eri, torch.eye(N, dtype=torch.float64, device=dev), h.contiguous(), nroe, 0.0, maxiter=2)
    print("ccsd", r["energies"], flush=True)
    del eri
    torch.cuda.empty_cache()
    # B formation at the (H2O)32 shape, FP32 GEMMs
    N, o, naux = 768, 160, 2432
    C, _ = torch.linalg.qr(torch.randn(N, N, dtype=torch.float64, device=dev, generator=g))
    A = torch.randn(naux, naux, dtype=torch.float64, device=dev, generator=g)
    J = (A @ A.T / naux + torch.eye(naux, dtype=torch.float64, device=dev)).contiguous()
    Pmn = torch.randn(naux, N, N, dtype=torch.float64, device=dev, generator=g)
    Pmn = (Pmn + Pmn.transpose(1, 2)).contiguous()
    M, _ = corr_amd.ri_metric(J, "cholesky")
    B = corr_amd.ri_form_bia(C.contiguous(), Pmn, M, o, "cholesky", "fp32")
    torch.cuda.synchronize()
    print("bform done", float(B.abs().max()), flush=True)


if __name__ == "__main__":
    main()
