"""Two-electron full CI in a spin-orbital basis (test oracle).

Determinants |p q>, p < q, over n spin orbitals; Slater rules with the
antisymmetrised integrals <pq||rs>:
  <pq|H|rs> = h_pr d_qs - h_ps d_qr - h_qr d_ps + h_qs d_pr + <pq||rs>.
For two electrons CCSD is exact, which pins oracle.ccsd.
"""
from __future__ import annotations

import numpy as np


def fci_two_electron(h: np.ndarray, V: np.ndarray) -> float:
    """Lowest electronic energy of two electrons (no nuclear repulsion)."""
    n = h.shape[0]
    dets = [(p, q) for p in range(n) for q in range(p + 1, n)]
    d = np.eye(n)
    Hm = np.empty((len(dets), len(dets)))
    for x, (p, q) in enumerate(dets):
        for y, (r, s) in enumerate(dets):
            Hm[x, y] = (h[p, r] * d[q, s] - h[p, s] * d[q, r] - h[q, r] * d[p, s] + h[q, s] * d[p, r]
                        + V[p, q, r, s])
    return float(np.linalg.eigvalsh(Hm)[0])
