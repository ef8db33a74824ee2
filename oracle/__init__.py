"""Oracle for the RMSD hot path — test infrastructure, never product code.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import, call, link or execute anything in this package, and
only as the checker: the product path (``mdtraj_amd``) never routes through
it and fails loudly when its HIP library is missing.

Contents
--------
``fp64``      numpy FP64 restatement of the reference mathematics (centering,
              traces, correlation matrix, QCP key matrix, largest eigenvalue,
              Kabsch rotation, superpose, no-superpose MSD).  This is the
              accuracy oracle (results within 1e-5 nm).
``rmsd32.c``  plain-C float32 restatement of the reference's own float32
              algorithm (double-accumulated centering, float M, float QCP
              coefficients, double root, float final step) with OpenMP over
              frames; it models the reference's float32 output and is the
              ``"port"`` CPU baseline.
``ref_driver.c`` + ``Makefile``
              compiles the reference's C++ core from its sources under
              ``/root/reference`` into ``oracle/_ref/`` (when present) so the
              restatements can be checked against the real thing and the CPU
              baseline can time it (``"reference"`` kind).

Parity status of each function is written in its docstring; every function
here is pinned by at least one check in ``tests/test_oracle*.py`` (golden
vectors from the compiled reference, closed forms, invariants, brute force).
"""
