"""CPU oracle for the msplat rasterizer path -- TEST INFRASTRUCTURE ONLY.

This package restates the reference algorithm (pointrix-project/msplat,
files cited per function) on the CPU: numpy for the per-Gaussian geometry,
the spherical harmonics and the sort; C (``blend.c``, built by ``build.py``)
for the per-pixel blending loops.  Every function takes ``dtype``:
``np.float32`` reproduces the GPU kernels' operation order (the geometry and
SH in bit-exact agreement with them), ``np.float64`` is the high-precision
reference the fp32 error is stated against.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker or the timed CPU
baseline; the product (``msplat_amd``) never imports or calls it.

Pinning (what checks each function, beyond the GPU comparison):
  geometry.project_*   closed-form pinhole with identity extrinsics; fp64 FD
  geometry.cov3d_*     R S^2 R^T closed form, identity quaternion, eigenvalues; fp64 FD
  geometry.ewa_*       isotropic closed form, brute-force integer rect; fp64 FD
  sh.*                 scipy.special.sph_harm (reference test mapping), golden
                       vectors of the reference polynomials (tests/golden), FD
  sort.*               reference KAT (test_sort_gaussian.py), brute force with
                       Python's stable sorted(), invariants
  blend.*              single-Gaussian closed form, empty tile, brute force on a
                       tiny image, fp64 FD of the adjoint
"""
