"""CPU oracle for the FE design-iteration hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import anything from this package, and only as the checker (or as the
timed CPU baseline).  The product path (``pymoto_amd``) never imports it and has no CPU fallback.

It restates, in numpy/scipy, the semantics of pyMOTO v2.0.1 for:

* grid numbering / connectivity        ``pymoto/common/domain.py:102-362``       (``oracle.fe``)
* element stiffness K0 (Gauss 2x2(x2))  ``pymoto/modules/assembly.py:318-463``    (``oracle.fe``)
* CSR pattern, datamap, np.add.at values and Dirichlet rows/cols
                                        ``pymoto/modules/assembly.py:38-296``     (``oracle.fe``)
* sensitivity contraction u_e^T K0 v_e  ``assembly.py:298-315``, ``dyadcarrier.py:286-414`` (``oracle.fe``)
* FilterConv forward/adjoint            ``pymoto/modules/filter.py:8-220``        (``oracle.filter``)
* SIMP MathExpression                   ``pymoto/modules/generic.py:14-140``      (``oracle.fe``)
* DampedJacobi / GeometricMultigrid / CG / LDAS / direct solve
                                        ``pymoto/solvers/iterative.py:21-403``,
                                        ``pymoto/solvers/solvers.py:6-306``,
                                        ``pymoto/solvers/sparse.py:522-550``     (``oracle.solvers``)
* one compliance design iteration       ``examples/topology_optimization/ex_compliance*.py``
                                                                                  (``oracle.pipeline``)

How each function is pinned (golden vectors from the reference, closed forms, brute force,
library routines) is listed in the header of each module and checked in ``tests/test_oracle*.py``.
"""
