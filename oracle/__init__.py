"""CPU oracle of the ideal-MHD step: TEST INFRASTRUCTURE ONLY.

A NumPy FP64 restatement of the reference's MHD step (carafe ``mhd.py``,
``mhd3d.py`` and the ``kernels/`` folders), written operation for operation in the
order of the reference's scalar (Pyccel) kernels so that an FP64 GPU path compiled
without FMA contraction can agree with it bit for bit.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker or the timed CPU
baseline.  The product (``wellbalanced_amd``) never imports it and has no CPU
fallback.

Pinning (what checks each function beyond retyping it):

* ``kernels.hlld_flux`` -- closed forms of the reference's tests
  (consistency, supersonic upwinding, mirror symmetry, isolated contact:
  ``tests/unit/test_mhd.py:38-74``), the HLLC limit at ``B = 0``, and golden
  vectors written by the reference's own ``mhd_sweep_pyccel.py``.
* ``kernels.sweep`` -- golden vectors of the reference's ``mhd_sweep_pyccel.py``
  on the seeded blocks of ``tests/unit/test_mhd_kernels.py:17-33`` (all limiters,
  both axes), plus the limiter closed forms of ``limiters.py:18-46``.
* ``kernels.ct_contact`` -- golden vectors of ``ct_contact_pyccel.py`` and the
  uniform-state reduction ``E_z = ecc``.
* ``kernels.primitive`` -- conserved -> primitive round trip from known primitive
  states, and the bad-cell counts of ``test_mhd_kernels.py:123-141``.
* ``kernels.rate`` -- closed forms of the fast speed (``B = 0``: the sound speed;
  field along the normal: ``max(a, b_n / sqrt(rho))``) and exact ``min rho``.
* ``mhd2d`` / ``mhd3d`` step glue -- invariants only (conservation, ``div B`` at
  round-off, uniform state, x/y symmetry, the z-uniform 3D run equal to the 2D
  run).  The flux-register update of the reference lives in the external
  ``patchwork`` package (not available); its operation order is this package's
  choice ``u - (dt/dx (F+ - F-) + dt/dy (G+ - G-))``: parity unpinned beyond the
  invariants.
"""
