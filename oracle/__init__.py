"""CPU oracle for the FEM elasticity hot path (test infrastructure only).

This package is the reference for correctness, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only to check or to time the CPU restatement beside the GPU
path.  Nothing in ``apple_amd`` imports it.

Contents
--------
``oracle.fem``      numpy FP64 restatement of the per-tet algebra, the kernel
                    templates (dV multiply, clamps, scatter), the model sum,
                    ``ExternalForce`` and the DOF map of liblaf/apple.
``oracle.pncg``     numpy FP64 restatement of the PNCG loop that apple's
                    benchmark states in-tree, plus a Newton polish and the
                    adjoint (sparse direct solves) used as end-state checks.
``oracle/c``        C (OpenMP) restatement of the same per-tet kernels with
                    separate passes and scatter, used as the timed CPU
                    baseline.

Each function's docstring cites the reference file:line it restates.  How
each function is pinned (closed forms, finite differences, the reference's
known-answer patch test and its volume invariant) is listed in
``tests/test_oracle_*.py``.
"""
