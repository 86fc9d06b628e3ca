"""CPU float64 oracle for the ceres_mi355x solver core.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (the ``ceres_mi355x``
package and its HIP library) imports, links or runs anything in this folder;
only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg use it, and only as the checker.

The oracle restates the reference's (``samionb/ceres-torch-cuda``) float64
trust-region semantics in numpy:

* ``lm.py``            the LM loop of ``src/ceres_torch/solver.py:35-357``
* ``ba.py``            Snavely BA evaluation + block-sparse Schur solves
* ``losses.py``        robust losses + Ceres corrector (``losses.py:13-193``)
* ``small_models.py``  curve fits and SE2 pose graphs
* ``generators.py``    the seeded synthetic workloads

How each function is pinned (golden vectors produced by running the reference
in the build container, reference known-answer tests, closed forms, finite
differences, brute force) is listed in the table at the top of
``tests/test_oracle.py``; every public function used as a checker appears
there with its pin.
"""
