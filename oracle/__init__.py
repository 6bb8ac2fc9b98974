"""CPU oracle for the DQMC sweep -- TEST INFRASTRUCTURE ONLY.

This package restates the reference's algorithm (lqmc/lqmc.py) in numpy so the
HIP path can be checked against it.  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it, and only as the
checker or as the timed CPU baseline.  The product package ``latticeqmc_amd``
never imports it and has no CPU fallback.

Modules and what pins each one (the tests named here hold the checks):
  model      K, E_K, lambda -- pinned by the reference's K fixtures
             (tests/golden/lattices.npz) and closed forms (test_oracle_model.py)
  mode_r     literal restatement of the reference fast sweep (lqmc.py:301-347)
             -- pinned bit-for-bit by the reference-generated trajectories
             tests/golden/mode_r_*.npz (test_oracle_mode_r.py)
  dqmc       textbook DQMC ("Mode D") with QR/UDV stabilisation, rank-1 and
             submatrix-delayed updates -- pinned by brute-force determinant
             ratios, inverse identities, U=0 closed forms, rank-1 vs delayed
             agreement and 2x2 physics vs exact (test_oracle_dqmc.py)
  det_mode   brute-force determinant sweep (lqmc.py:236-299) -- pinned by
             agreement of its accept bits with dqmc on the same uniforms
  exact      Fock-space exact diagonalisation and Trotterised exact values --
             pinned by U=0 free-fermion formulas and known 2-site results
  philox     Philox4x32-10 counter RNG -- pinned by the Random123 known-answer
             vectors (test_oracle_philox.py)
"""
