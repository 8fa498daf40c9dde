"""Readers and writers for the reference's binary input files.

All files are raw native-endian (little-endian x86-64) records with no headers
beyond the ones listed.  Each layout follows the reference's reader and the
psi4 plugin's writer:

  PREFIX.sys    read_system (src/IO/ops_io.cc:63-111), plugin.cc:637-653
                int32 nroe, nroa, nroao, naux_1, naux_2;
                int64 nrofint, nrofaux, nrofaux2;
                float64 coord[3 nroa], charges[nroa], zeff[nroa], mass[nroa];
                basis names (the reader takes three 32-byte fields)
  PREFIX.oei    read_oei (ops_io.cc:114-124), plugin.cc:252-256
                float64 H, T, S, V, each nroao x nroao row-major
  PREFIX.tei    read_tei (ops_io.cc:203-212), plugin.cc:329-332
                int64 sortcount[4]; float64 intval[nrofint]; uint16 intnums[4 nrofint]
  PREFIX.ahfw   read_wav_HF (ops_io.cc:283-299), write_wav_HF (:265-273)
  PREFIX.bhfw   plugin.cc:559-569 (beta orbitals, same layout)
                int32 nroao; float64 MOens[nroao]; float64 MOs[nroao x nroao] (row = MO)
  PREFIX.J      plugin.cc:425-434: float64 J[naux x naux], the AO metric (P|Q)
  PREFIX.rimp2  read_transform_ri (src/POSTHF/ri.cc:109-128), plugin.cc:518-521:
                float64 BPQ[naux x nroao x nroao] = J^-1/2 (P|mu nu)

Large arrays (.tei, .rimp2) are memory-mapped, not read whole.  Files too short
for the sizes they declare raise ValueError, where the reference would read
past the end.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

_I32 = np.dtype("<i4")
_I64 = np.dtype("<i8")
_F64 = np.dtype("<f8")
_U16 = np.dtype("<u2")
NAME_BYTES = 32   # each basis name field of the .sys reader (ops_io.cc:68-70, 93-95)


@dataclass
class SystemInfo:
    """The fields of systeminfo that PREFIX.sys carries (src/UTIL/util.h:16-45)."""
    nroe: int
    nroa: int
    nroao: int
    naux_1: int
    naux_2: int
    nrofint: int
    nrofaux: int
    nrofaux2: int
    coord: np.ndarray            # (nroa, 3) bohr
    charges: np.ndarray          # (nroa,)
    zeff: np.ndarray             # (nroa,)
    mass: np.ndarray             # (nroa,)
    basis_names: tuple = field(default_factory=lambda: ("", "", ""))   # OB, JK, RI

    @property
    def ion_rep(self) -> float:
        """Nuclear repulsion sum_{A<B} Z_A Z_B / R_AB (calc_ion_rep, ops_rhf.cc:283-292)."""
        e = 0.0
        for a in range(self.nroa):
            for b in range(a + 1, self.nroa):
                e += self.charges[a] * self.charges[b] / float(np.linalg.norm(self.coord[a] - self.coord[b]))
        return e


def _need(buf_len: int, need: int, path: str) -> None:
    if buf_len < need:
        raise ValueError(f"{path}: {buf_len} bytes, the declared sizes need {need}")


def _name_field(raw: bytes) -> str:
    """The reader streams a 32-byte field into a std::string with >>, i.e. the
    first whitespace-delimited token (a NUL also ends it here)."""
    text = raw.split(b"\0", 1)[0].decode("latin-1")
    parts = text.split()
    return parts[0] if parts else ""


def read_system(path: str) -> SystemInfo:
    """PREFIX.sys (read_system, ops_io.cc:63-111)."""
    buf = open(path, "rb").read()
    head = 5 * 4 + 3 * 8
    _need(len(buf), head, path)
    nroe, nroa, nroao, naux_1, naux_2 = (int(x) for x in np.frombuffer(buf, _I32, 5, 0))
    nrofint, nrofaux, nrofaux2 = (int(x) for x in np.frombuffer(buf, _I64, 3, 20))
    if nroa < 0 or nroao < 0:
        raise ValueError(f"{path}: negative sizes")
    arrays = 6 * nroa * 8
    _need(len(buf), head + arrays, path)
    f = np.frombuffer(buf, _F64, 6 * nroa, head)
    coord = f[: 3 * nroa].reshape(nroa, 3).copy()
    charges, zeff, mass = (f[3 * nroa + k * nroa: 4 * nroa + k * nroa].copy() for k in range(3))
    tail = buf[head + arrays:]
    # The plugin resizes each name to 32 bytes, NUL-padded (plugin.cc:596-598,
    # 650-652); a file that ends early yields empty names.
    names = tuple(_name_field(tail[k * NAME_BYTES:(k + 1) * NAME_BYTES]) for k in range(3))
    return SystemInfo(nroe, nroa, nroao, naux_1, naux_2, nrofint, nrofaux, nrofaux2, coord, charges, zeff, mass,
                      names)


def write_system(path: str, s: SystemInfo) -> None:
    """PREFIX.sys in the reader's layout; each basis name is NUL-padded to 32 bytes."""
    with open(path, "wb") as fh:
        fh.write(np.array([s.nroe, s.nroa, s.nroao, s.naux_1, s.naux_2], _I32).tobytes())
        fh.write(np.array([s.nrofint, s.nrofaux, s.nrofaux2], _I64).tobytes())
        fh.write(np.ascontiguousarray(s.coord, _F64).reshape(-1).tobytes())
        for a in (s.charges, s.zeff, s.mass):
            fh.write(np.ascontiguousarray(a, _F64).tobytes())
        for name in s.basis_names:
            raw = name.encode("latin-1")
            if len(raw) >= NAME_BYTES:
                raise ValueError(f"basis name {name!r} longer than {NAME_BYTES - 1} bytes")
            fh.write(raw + b"\0" * (NAME_BYTES - len(raw)))


def read_oei(path: str, nroao: int) -> dict:
    """PREFIX.oei (read_oei, ops_io.cc:114-124): H, T, S, V."""
    n2 = nroao * nroao
    a = np.fromfile(path, _F64)
    _need(a.size * 8, 4 * n2 * 8, path)
    return {k: a[i * n2:(i + 1) * n2].reshape(nroao, nroao).copy() for i, k in enumerate(("H", "T", "S", "V"))}


def write_oei(path: str, H, T, S, V) -> None:
    with open(path, "wb") as fh:
        for m in (H, T, S, V):
            fh.write(np.ascontiguousarray(m, _F64).tobytes())


def read_tei(path: str, nrofint: int):
    """PREFIX.tei (read_tei, ops_io.cc:203-212): (sortcount[4], intval, intnums[n, 4]).
    intval and intnums are read-only memory maps of the file."""
    need = 4 * 8 + nrofint * 8 + nrofint * 4 * 2
    _need(os.path.getsize(path), need, path)
    sortcount = np.fromfile(path, _I64, 4).copy()
    intval = np.memmap(path, _F64, "r", offset=32, shape=(nrofint,))
    intnums = np.memmap(path, _U16, "r", offset=32 + nrofint * 8, shape=(nrofint, 4))
    if not (0 <= sortcount[0] <= sortcount[1] <= sortcount[2] <= sortcount[3] <= nrofint):
        raise ValueError(f"{path}: sortcount {sortcount.tolist()} is not non-decreasing within {nrofint}")
    return sortcount, intval, intnums


def write_tei(path: str, sortcount, intval, intnums) -> None:
    with open(path, "wb") as fh:
        fh.write(np.asarray(sortcount, _I64).reshape(4).tobytes())
        fh.write(np.ascontiguousarray(intval, _F64).tobytes())
        fh.write(np.ascontiguousarray(intnums, _U16).reshape(-1).tobytes())


def read_wav(path: str, nroao: int | None = None):
    """PREFIX.ahfw / .bhfw (read_wav_HF, ops_io.cc:283-299): (MOens, MOs) with
    MOs[i, mu] (row = MO).  A stored size other than nroao is refused, as the
    reference does ("Wrong HF wavefunction size")."""
    buf = open(path, "rb").read()
    _need(len(buf), 4, path)
    n = int(np.frombuffer(buf, _I32, 1, 0)[0])
    if nroao is not None and n != nroao:
        raise ValueError(f"{path}: wrong HF wavefunction size {n} (expected {nroao})")
    _need(len(buf), 4 + (n + n * n) * 8, path)
    ens = np.frombuffer(buf, _F64, n, 4).copy()
    mos = np.frombuffer(buf, _F64, n * n, 4 + 8 * n).reshape(n, n).copy()
    return ens, mos


def write_wav(path: str, MOens, MOs) -> None:
    """write_wav_HF (ops_io.cc:265-273)."""
    MOs = np.ascontiguousarray(MOs, _F64)
    n = MOs.shape[0]
    with open(path, "wb") as fh:
        fh.write(np.array([n], _I32).tobytes())
        fh.write(np.ascontiguousarray(MOens, _F64).reshape(n).tobytes())
        fh.write(MOs.tobytes())


def read_metric(path: str, naux: int) -> np.ndarray:
    """PREFIX.J (plugin.cc:425-434): the AO metric (P|Q), naux x naux."""
    _need(os.path.getsize(path), naux * naux * 8, path)
    return np.fromfile(path, _F64, naux * naux).reshape(naux, naux)


def write_metric(path: str, J) -> None:
    with open(path, "wb") as fh:
        fh.write(np.ascontiguousarray(J, _F64).tobytes())


def read_bpq(path: str, naux: int, nroao: int) -> np.ndarray:
    """PREFIX.rimp2 (read_transform_ri, ri.cc:125-128): BPQ[Q, mu, nu], memory-mapped."""
    _need(os.path.getsize(path), naux * nroao * nroao * 8, path)
    return np.memmap(path, _F64, "r", shape=(naux, nroao, nroao))


def write_bpq(path: str, BPQ) -> None:
    with open(path, "wb") as fh:
        fh.write(np.ascontiguousarray(BPQ, _F64).tobytes())
