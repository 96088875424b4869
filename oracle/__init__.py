"""CPU FP64 oracle for the Green--Laplace bound-wave hot path.

TEST INFRASTRUCTURE ONLY.  This package restates the reference algorithm in
NumPy (complex128, ``numpy.fft``) so that the HIP path can be checked against
it.  Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg
of ``bench.py`` may import it, and only as the checker or as the timed CPU
baseline.  The product package never imports it and has no CPU fallback.

Pinning (what checks each function besides re-evaluating its own formula):

* ``gauss_laguerre_rule``: exact factorial moments, the frozen GL8 nodes and
  weights of the reference interface JSON, and the closed-form psi22 two-node
  rule.
* ``reconstruct_directional_field`` / ``gl_spectral_surface`` orders 1-3: the
  reference's three MATLAB-generated 64x64 fixtures at the reference's own
  tolerances, the linear closed form, exact homogeneity of degree 2 and 3, and
  (eta22) the ordered-pair brute force of the same GL kernel.
* ``wit`` (``oracle/wit_oracle.cpp``, sectors 44, 40 and 42): the reference
  executable built from its own source into ``oracle/_ref`` (n0050, n0100,
  n0382 inputs), the reference's shipped n0382 output, and the reference
  executable's n0050 / n0200 outputs of the signed sectors; every counter
  exact.
* ``modal_oracle`` (single-direction time path): the reference's two
  MATLAB-generated time records (eta22/eta33 at 3e-13, eta20 at 8e-11), the
  on-grid identity with the spatial graph at 3e-14, the off-grid ladder
  convergence, degree-2/3 homogeneity, and the GL node sums against the
  Laplace integrals they approximate.
"""
