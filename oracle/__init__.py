"""CPU oracle for the electron-repulsion-integral path.  TEST INFRASTRUCTURE ONLY.

This package restates, in NumPy, the reference's Obara-Saika / Head-Gordon-Pople
ERI algorithm (gbasis/integrals/_two_elec_int_improved.py:591-732), its Boys
function (gbasis/integrals/boys_functions.py:26-63), its Schwarz screening
(gbasis/integrals/_schwarz_screening.py:14-90,169-191) and its four-index
assembly (gbasis/base_four_symm.py:160-389, gbasis/integrals/electron_repulsion.py:214-504).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker.  The product
(``gbasis_amd``) never imports it and never falls back to it.

How each piece is pinned (so that a dropped term, a wrong sign or index, or a
transposed operand fails a test):
  * boys.py        closed forms F_m(0)=1/(2m+1), F_0 via erf, the downward
                   recurrence, and mpmath at 50 digits.
  * os_hgp.py      the reference test-suite's HORTON primitive values for s..f
                   components at a fixed geometry, the closed-form (ss|ss), an
                   independent McMurchie-Davidson evaluation in mpmath at random
                   geometries (components up to d), and the 8-fold
                   permutational symmetry.
  * erf / erfc     (boys_potential) the closed form of Gaussian charge clouds
                   screened by a Gaussian of exponent omega^2, and its centre
                   derivative for a p function.
  * assembly.py    the reference fixture data_horton_bec_cart_elec_repulsion.npy
                   (Be+C STO-6G, physicist notation), the Schwarz inequality, and
                   vectors written by the reference itself
                   (tests/golden/reference_vectors.npz): its Schwarz matrices
                   (identical screened lists) and sampled dense FP64 ERIs, within
                   the reference's own FP64 error (tests/test_reference_vectors.py).
  * spherical tables (product module gbasis_amd.spherical, used here) are pinned
                   by HORTON cartpure matrices (tests/test_spherical.py).
The tests of boys.py, os_hgp.py and assembly.py are in tests/test_oracle_golden.py.
"""
