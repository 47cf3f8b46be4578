"""CPU FP64 restatement of Updes' dense RBF-collocation path -- TEST INFRASTRUCTURE ONLY.

This package is the oracle: a plain NumPy/SciPy restatement of the reference
algorithm (ddrous/Updes v1.0.3, ``updes/{utils,cloud,assembly,operators}.py``)
used to check the HIP path.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only to check results
or as the timed CPU baseline.  The product package ``updes_amd`` never imports it
and has no CPU fallback.

Modules
-------
rbf       closed forms of the kernels, their derivatives and the monomial basis
          (reference ``updes/utils.py:19-144`` and the autodiff operators in
          ``updes/operators.py:15-111``).
cloud     ``SquareCloud`` numbering / node types / normals / renumbering
          (reference ``updes/cloud.py:10-172, 378-510``).
assembly  Phi, P, A, op(Phi,P), bd(Phi,P), B and q exactly as the reference
          assembles them (reference ``updes/assembly.py``).
solver    the reference solve (``inv(A)``, ``D @ inv(A)``, QR) and the
          equivalent augmented-system solve, plus field evaluation
          (reference ``updes/operators.py:118-371, 512-683``).

Pinning (what makes the restatement trustworthy without running JAX):
* every closed form is checked against SymPy derivatives evaluated in mpmath
  at 50 digits (``tests/test_oracle_pins.py``);
* polynomial reproduction, Dirichlet-row identity of B, value-at-node identity,
  the integral of x^2/(1+y^2) over the unit square (pi/12) and the renumbering
  permutation are checked as invariants;
* a brute-force mpmath solve pins the linear algebra on tiny clouds;
* the committed golden vectors under ``tests/golden`` freeze the oracle's
  outputs; planted mistakes in the oracle are shown to fail the pins.
The reference's own tests hold no golden vectors for ``pde_solver`` (their
tightest tolerance on the solve is 1e-2), so no part of the solve can be pinned
against reference-produced numbers; where the pins above do not reach, the
docstrings say "parity unpinned".
"""
