"""CPU FP64 oracle for the FEA forward + adjoint path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker.  The product path (``jaxsso_amd``) never imports it and fails loudly
when its HIP library is missing.

It is a numpy/scipy restatement of the reference algorithm (GaoyuanWu/JaxSSO):

* ``elements``  -- beam-column and MITC-4 element stiffness
  (``JaxSSO/element.py:66-149`` and ``:471-1106``), written so the same code
  runs in float64 and in complex128 (complex-step derivatives).
* ``system``    -- COO assembly with duplicate summation, the scalar sparsity
  pattern, the Lagrange-augmented ``K_aug``/``f_aug``
  (``JaxSSO/assemblemodel.py:111-236``).
* ``solve``     -- the reference's SuperLU solve on ``K_aug``
  (``JaxSSO/solver.py:176-218``) and an equilibrated reduced-system solve with
  extended-precision iterative refinement (the accurate FP64 oracle).
* ``gradient``  -- compliance ``C = 1/2 f.u`` and its gradient
  ``-lambda_e^T dK_e/dp u_e`` (``JaxSSO/solver.py:221-250`` with
  ``JaxSSO/SSO_model.py:297-339``), element derivatives by complex step.

How each part is pinned (tests/test_oracle*.py): the reference notebooks'
recorded values (arch, barrel), closed forms (cantilever, axial, torsion),
invariants (symmetry, rigid-body null space, patch tests, linearity in E,
t^3 bending scaling, homogeneity, translation invariance of the gradient),
brute-force dense assembly and central finite differences.
"""
