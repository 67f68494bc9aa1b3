"""
CPU FP64 oracle of the GFN1/GFN2-xTB single-point path (TEST INFRASTRUCTURE).

This package is a plain-torch restatement of the reference algorithms
(marvinfriede/dxtb 0.4.0), written to check the HIP path.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker or the timed CPU baseline -- never as the
thing measured or shipped.  The product path (``dxtb_amd``) never imports it.

Every function cites the reference file:line it follows.  How each part is
pinned (reference fixtures, closed forms, brute force) is recorded in the
test that pins it:

* ``basis``      -- STO-nG expansion, H-2s orthogonalisation, AO ordering.
                    Pinned by tests/test_oracle_integrals.py (unit self
                    overlap, orthogonality, closed forms).
* ``integrals``  -- overlap / dipole / quadrupole via the Obara-Saika
                    three-index recursion with origin 0, mirror, snorm,
                    r0->rj shifts, traceless.  Pinned by Gauss-Hermite
                    quadrature, one-centre closed forms and the reference's
                    float32 overlap fixtures.
* ``hamiltonian``-- CN (D3 type, exp/gfn2 counting), H0.  Pinned by the
                    reference's tblite H0 fixtures and dCN fixtures.
* ``coulomb``    -- ES2, ES3, AES2 matrices/energies/potentials, repulsion.
                    Pinned by the reference's coulomb/aes2/repulsion fixtures.
* ``scf``        -- Fermi smearing, density, charges, potentials, Fock
                    build, generalized eigensolve, Broyden fixed point,
                    free energy, variational forces.  Pinned by the
                    reference's GFN1 tblite SCF energies and by central finite
                    differences of the energy.
"""
