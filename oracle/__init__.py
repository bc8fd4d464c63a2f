"""FP64 CPU oracle for the MHM local-problem stage.

TEST INFRASTRUCTURE ONLY.  This package restates, in NumPy/SciPy, the
reference algorithm of PyMHM's local stage (local meshes, P1/P2 assembly,
trace coupling, the ``LocalProblem`` contract, equilibrated SuperLU
condensation, condensed contributions, ordered scatter and reconstruction).
It is the checker that the HIP path is compared against.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it; the product package ``pymhm_amd`` never does.

Each function cites the reference file:line it follows (paths relative to the
reference repository root).  How each part is pinned:

* meshes / P2 node order / trace coupling / compensated assembly /
  condensation / scatter / reconstruction: the reference's golden vectors
  ``benchmarks/results/numba-20261009/original-2d-2d-n4-r8-p2-q6.npz`` and
  ``...-3d-3d-n2-r4-p2-q6.npz`` (copied under ``tests/golden/``), checked in
  ``tests/test_oracle_golden.py``;
* P1 element operators: closed forms (``tests/test_oracle_algebra.py``);
* discontinuous-P1 Legendre face modes, their orientation on the second owner
  of a face and the degree-1 trace coupling: ``legvander``, the segment mass
  L/(2j+1), hand-derived closed forms of int phi_i L_j on every fine edge, and
  the affine patch test with weak Dirichlet data (``boundary_data``) on square
  and Voronoi macros (``tests/test_oracle_pins.py``);
* ``PolygonMesh`` ear split and refinement: area, kept vertices, orientation
  and conformity on the reference's L-shaped macro and on Voronoi macros, and
  the patch test on Voronoi macros (``tests/test_oracle_pins.py``);
* condensation algebra: known answers of ``tests/test_hybrid_contributions.py``
  and ``tests/test_hybrid.py`` of the reference, re-derived in
  ``tests/test_oracle_algebra.py``, plus a monolithic dense solve;
* SuperLU oracle semantics: cross-checked with a dense LAPACK solve.
"""
