"""CPU FP64 oracle for the DWBA radial multipole double integrals I_L.

TEST INFRASTRUCTURE ONLY.  This package restates, in numpy (long double for
the O(N) contraction, mpmath for the brute force), the reference algorithm of
StaryPasterz/DWBA's ``radial_ME_all_L`` (dwba_matrix_elements.py:780-1300) and
the Filon / Clenshaw-Curtis machinery it calls (oscillatory_integrals.py).  It
is the checker of the HIP product path in ``dwba_amd``: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only to check or to time a CPU baseline.  The product package
never imports it.

Modules
-------
grid        grid.py:419-611 (trapezoid / non-uniform Simpson weights, exp grid)
quadrature  oscillatory_integrals.py:120-302, 1456-1755 (CC weights, phase
            nodes, adaptive phase increment, log-r interpolation)
tail        oscillatory_integrals.py:312-535, 957-1451, 2064-2184 (outer tail:
            sinA*sinB phases, Filon, Levin, selector, analytical tail)
radial      dwba_matrix_elements.py:780-1300 + oscillatory_integrals.py:2385-2780
            (idx_limit, the faithful dense restatement that materialises K_L,
            and the alpha/beta reduction I_L = alpha^T K_L beta)
semisep     exact O(N) prefix-sum form of alpha^T K_L beta in long double
brute       mpmath brute force for tiny N
gradient    d/d rho_a, d/d rho_b of sum_L c_L I_L (K_L x by the same prefix sums)

The seeded synthetic inputs of the benchmark workloads live in the product
package (dwba_amd/synthetic.py): they are inputs, not a checker.

How each function is pinned (checks live in tests/test_oracle_*.py):
  grid.simpson_weights     scipy.integrate.simpson (odd N, same non-uniform
                           formula); closed form of the trailing trapezoid
  quadrature.cc_ref        closed-form weight tables; polynomial exactness of
                           clenshaw_curtis_nodes; the 2.0667 weight sum of the
                           5-node module formula
  quadrature.phase_nodes   closed-form node count (eta=0); the stepping
                           invariant dr = dphi/|k+eta/r| (eta!=0)
  quadrature.interp        exactness on functions linear in log r (log grids)
                           and linear in r (other grids); np.interp special case
  radial (reduced)         against radial (faithful dense) on the same inputs:
                           two independent routes to the same number
  semisep                  mpmath brute force at N<=48; hydrogenic closed forms
                           F0(1s,1s)=5/8, G1(1s,2p)=112/2187; factorisation
                           when the supports are disjoint; symmetry
  tail.filon_segment       exactness of the Taylor branch on quadratic
                           envelopes with linear phase (textbook Filon)
  tail.levin_segment       exactness on f = -p' + i w p, p of degree <= 7, linear
                           phase (the collocation system with the reference's
                           descending-point Chebyshev matrix on ascending nodes)
  tail.filon_integral      the vectorised segment sweep equals the loop over
                           filon_segment; quad branch: closed form within quad's
                           tolerance
  tail.outer_integral,     golden vectors written by tests/golden/make_golden.py
  radial (composite, H     (which imports the reference's CPU modules where they
  1s->2p at 50 eV)         are importable; the fixtures travel, the reference does not)
"""
