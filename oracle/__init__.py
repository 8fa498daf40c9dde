"""CPU FP64 oracle for the corr post-HF contraction path.

TEST INFRASTRUCTURE ONLY.  This package restates the reference's arithmetic
(Masterluke87/corr, src/POSTHF) in numpy so that the HIP path can be checked
against it.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker or
as the timed CPU baseline.  The product path (``corr_amd``) never calls it.

Modules
  posthf  build_FMo, the sorted unique ERI list (resort_integrals), its
          unfold (trans_I), the 4-index transform (mo2int_op) and its
          brute-force spec (calc_mo2int_bf), canonical MP2, the RI metric
          (calculate_ri), transform_ri, run_canonical_mp2_ri and the RI-MP2
          energy gradient.
  ccsd    spin-orbital CCSD of ccsd_ur with every intermediate.
  fci     two-electron full CI (pins CCSD: exact for two electrons).

Parity pins for every function are listed in each module's docstring.
"""
