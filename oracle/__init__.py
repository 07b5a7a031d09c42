"""CPU FP64 oracle for the PARSEC.py SCF hot path -- TEST INFRASTRUCTURE ONLY.

This package is a from-scratch NumPy/SciPy restatement of the reference
algorithm (zeyizhang-berkeley/parsec_python).  Each function cites the
reference file:line it follows.  It exists to check the HIP product path:

* only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
  ``bench.py`` may import it, and only as the checker / the timed CPU baseline;
* the product package ``parsec_python_amd`` never imports it and has no CPU
  fallback through it.

Pinning status of each function (what checks it besides its own formula):

==========================  ==================================================
function                    pinned by
==========================  ==================================================
grid.size_domain            reference logs: nx 52/92/22, rad 9.63760/17.19651/6.3
fd.fd_coefficients          closed form + Taylor test (orders 2..12);
                            hard-coded reference fractions (orders 2/4/6/8)
fd.laplacian                brute-force dense stencil on tiny grids
splines.*                   scipy CubicSpline(bc_type='natural') + clamping
pseudo.pseudo_diag          H2O and H golden SAD densities (bit-level ~1e-13),
                            logged electron counts and sum(rho0*V_H0)
pseudo.pseudo_nl            brute-force per-point loop restatement on a small
                            grid; symmetric PSD-sign structure; CSR == factored
pseudo.nuclear              logged E_nn (H2O 13.76119 Ry, C10H8 454.93231 Ry)
xc.exc_nspn                 logged initial E_xc (H2O, C10H8, H); PZ81 closed
                            form continuity at rs = 1
eig.* (filters, RR, bounds) Chebyshev identity on diagonal H; Ritz values vs
                            scipy eigsh on small H
eig.occupations             closed-form Fermi-Dirac; electron count invariant
pcg.pcg                     direct sparse solve; reference first-step aliasing
mixer.Msecant1              Broyden-I identity: secant equation on a linear
                            model; restart rule
scf.run_scf                 H2O sample log: E_total -34.20958 Ry, occupied
                            eigenvalues, converged density fixture
==========================  ==================================================
"""
