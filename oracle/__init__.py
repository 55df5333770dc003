"""CPU oracle for the PM + COLA hot path -- TEST INFRASTRUCTURE ONLY.

This package restates, in FP64 numpy/scipy, the algorithms of the fs2
reference that the HIP path replaces, to check it against: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
it, and never as the thing measured or shipped.  ``fs2_amd`` never imports it.

Modules and what pins each of them (checks live in ``tests/test_oracle*.py``):

* ``cosmology`` -- py/cosmology.cpp:24-171.  Pinned by the hypergeometric
  closed form of D1 and by the EdS limit (Omega_m = 1: D1 = a, f = 1).
* ``cic``       -- py/pm.cpp:57-220.  Pinned by the one-particle known answers
  of test/test_pm_cic.py:46-51, by closed-form CIC weights on hand-placed
  particles, by mass conservation and by the paint/readout adjoint identity.
* ``pm``        -- py/pm.cpp:306-508.  Pinned by the plane-wave closed form
  (delta = A cos(k.x) gives f = -A k/|k|^2 sin(k.x) on the mesh) and by
  momentum conservation.
* ``cola``      -- py/cola.cpp:48-179.  Pinned by the EdS closed forms of the
  kick factor and of the drift integral.
* ``lpt``       -- py/lpt.cpp:100-225, 439-591.  Pinned by div Psi1 = -delta,
  by the closed-form 2LPT source of a single plane wave (zero) and of two
  crossed plane waves, and by the lattice/id layout of
  test/test_particles_h5.py:57-77.
* ``power``     -- py/power.cpp:15-171 (table + natural cubic spline) and
  Eisenstein & Hu 1998 no-wiggle transfer function (eqs. 26-31).  Pinned by
  the spline reproducing the table nodes and cubic polynomials exactly, and
  by sigma_8 normalisation.
* ``rng``       -- Philox4x32-10 (Salmon et al., SC'11).  Pinned by the
  published known-answer vectors of the Random123 distribution.
* ``pk``        -- binned power-spectrum estimator.  Pinned by a single
  plane wave of known amplitude.
"""
