"""Readers/writers of the reference's input files (corr_amd/formats.py).

Byte layouts are checked against files assembled field by field in the order
the reference's writers emit them (plugin.cc:252-256, 329-332, 425-434,
518-521, 559-569, 637-653; write_wav_HF ops_io.cc:265-273), then every format
is round-tripped.  No GPU needed."""
import struct

import numpy as np
import pytest

from corr_amd import formats as FM
from oracle import posthf as O
from oracle import scf as SCF
from tests import _synth as S


def _sys_bytes(nroe, nroao, naux1, naux2, nrofint, nrofaux, nrofaux2, coord, charges, zeff, mass, names):
    nroa = len(charges)
    b = struct.pack("<5i", nroe, nroa, nroao, naux1, naux2) + struct.pack("<3q", nrofint, nrofaux, nrofaux2)
    b += struct.pack(f"<{3 * nroa}d", *np.ravel(coord))
    for a in (charges, zeff, mass):
        b += struct.pack(f"<{nroa}d", *a)
    for n in names:                      # std::string::resize(32): NUL padded
        b += n.encode() + b"\0" * (32 - len(n))
    return b


def test_read_system_reference_layout(tmp_path):
    coord = np.array([[0.0, 0.0, 0.0], [0.0, 1.43, 1.1], [0.0, -1.43, 1.1]])
    charges, zeff, mass = np.array([8.0, 1, 1]), np.array([8.0, 1, 1]), np.array([16.0, 1, 1])
    raw = _sys_bytes(10, 24, 116, 84, 12345, 0, 0, coord, charges, zeff, mass, ["DEF2-SVP", "DEF2-SVP-JKFIT",
                                                                                 "DEF2-SVP-RI"])
    p = tmp_path / "w.sys"
    p.write_bytes(raw)
    s = FM.read_system(str(p))
    assert (s.nroe, s.nroa, s.nroao, s.naux_1, s.naux_2) == (10, 3, 24, 116, 84)
    assert (s.nrofint, s.nrofaux, s.nrofaux2) == (12345, 0, 0)
    assert np.array_equal(s.coord, coord) and np.array_equal(s.charges, charges)
    assert np.array_equal(s.zeff, zeff) and np.array_equal(s.mass, mass)
    assert s.basis_names == ("DEF2-SVP", "DEF2-SVP-JKFIT", "DEF2-SVP-RI")
    assert abs(s.ion_rep - SCF.calc_ion_rep(coord, charges)) < 1e-15
    q = tmp_path / "w2.sys"
    FM.write_system(str(q), s)
    assert q.read_bytes() == raw


def test_system_name_token_and_short_file(tmp_path):
    raw = _sys_bytes(2, 2, 2, 0, 0, 0, 0, np.zeros((2, 3)), np.ones(2), np.ones(2), np.ones(2),
                     ["sto-3g  trailing", "", "x"])
    p = tmp_path / "a.sys"
    p.write_bytes(raw)
    assert FM.read_system(str(p)).basis_names == ("sto-3g", "", "x")   # stream >> takes one token
    p.write_bytes(raw[:40])
    with pytest.raises(ValueError):
        FM.read_system(str(p))


def test_oei_roundtrip_and_order(tmp_path):
    rng = np.random.default_rng(0)
    mats = [rng.standard_normal((5, 5)) for _ in range(4)]
    p = tmp_path / "a.oei"
    p.write_bytes(b"".join(m.astype("<f8").tobytes() for m in mats))   # H, T, S, V
    o = FM.read_oei(str(p), 5)
    for k, m in zip("HTSV", mats):
        assert np.array_equal(o[k], m)
    FM.write_oei(str(tmp_path / "b.oei"), *mats)
    assert (tmp_path / "b.oei").read_bytes() == p.read_bytes()
    with pytest.raises(ValueError):
        FM.read_oei(str(p), 6)


def test_tei_roundtrip(tmp_path):
    dense = S.eri_like(5, seed=2)
    val, idx, sc = O.sorted_unique_list(dense)
    p = tmp_path / "a.tei"
    p.write_bytes(np.asarray(sc, "<i8").tobytes() + val.astype("<f8").tobytes() + idx.astype("<u2").tobytes())
    sc2, v2, i2 = FM.read_tei(str(p), len(val))
    assert np.array_equal(sc2, sc) and np.array_equal(v2, val) and np.array_equal(i2, idx)
    assert np.array_equal(O.unfold_sorted(5, v2, i2, sc2), dense)
    FM.write_tei(str(tmp_path / "b.tei"), sc, val, idx)
    assert (tmp_path / "b.tei").read_bytes() == p.read_bytes()
    with pytest.raises(ValueError):
        FM.read_tei(str(p), len(val) + 1)
    bad = np.array([3, 2, 5, 7], "<i8")
    (tmp_path / "c.tei").write_bytes(bad.tobytes() + val.astype("<f8").tobytes() + idx.astype("<u2").tobytes())
    with pytest.raises(ValueError):
        FM.read_tei(str(tmp_path / "c.tei"), len(val))


def test_wav_roundtrip_and_size_check(tmp_path):
    ens = np.linspace(-1, 1, 4)
    C = S.orthonormal(4, seed=3)
    p = tmp_path / "a.ahfw"
    p.write_bytes(struct.pack("<i", 4) + ens.astype("<f8").tobytes() + C.astype("<f8").tobytes())
    e2, C2 = FM.read_wav(str(p), 4)
    assert np.array_equal(e2, ens) and np.array_equal(C2, C)
    FM.write_wav(str(tmp_path / "b.ahfw"), ens, C)
    assert (tmp_path / "b.ahfw").read_bytes() == p.read_bytes()
    with pytest.raises(ValueError, match="wrong HF wavefunction size"):
        FM.read_wav(str(p), 5)


def test_metric_and_bpq_roundtrip(tmp_path):
    rng = np.random.default_rng(4)
    J = S.metric(6, seed=1)
    B = rng.standard_normal((6, 3, 3))
    FM.write_metric(str(tmp_path / "a.J"), J)
    FM.write_bpq(str(tmp_path / "a.rimp2"), B)
    assert (tmp_path / "a.J").stat().st_size == 6 * 6 * 8
    assert np.array_equal(FM.read_metric(str(tmp_path / "a.J"), 6), J)
    assert np.array_equal(FM.read_bpq(str(tmp_path / "a.rimp2"), 6, 3), B)
    with pytest.raises(ValueError):
        FM.read_bpq(str(tmp_path / "a.rimp2"), 7, 3)


def test_sorted_unique_list_torch_matches_oracle():
    """corr_amd.pipeline.sorted_unique_list (torch, any device) against the
    oracle's restatement of calculate_libint_tei's selection + resort_integrals
    (ops_rhf.cc:682-693, 760-896): the same class boundaries and, within each
    class, the same (a, b, c, d, value) entries; the unfolded tensor equals
    the input."""
    import torch
    from corr_amd.pipeline import sorted_unique_list
    for N, seed in ((5, 2), (7, 9)):
        dense = S.eri_like(N, seed=seed)
        val, idx, sc = O.sorted_unique_list(dense, thresh=-1.0)
        tv, ti, tsc = sorted_unique_list(torch.from_numpy(dense))
        assert list(tsc) == [int(x) for x in sc]
        tv, ti = tv.numpy(), ti.numpy().astype(np.int64)
        bounds = [0, *sc, len(val)]
        for k in range(5):
            lo, hi = bounds[k], bounds[k + 1]
            mine = sorted(zip(map(tuple, ti[lo:hi]), tv[lo:hi]))
            ref = sorted(zip(map(tuple, idx[lo:hi].astype(np.int64)), val[lo:hi]))
            assert mine == ref
        assert np.array_equal(O.unfold_sorted(N, tv, ti.astype(np.uint16), np.array(tsc)), dense)
