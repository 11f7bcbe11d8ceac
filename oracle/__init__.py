"""CPU oracle for the jaccpot FMM evaluation path (test infrastructure only).

This package is a float64 NumPy restatement of the algorithm that jaccpot runs
on its real (Dehnen no-sqrt2) basis path, written from the reference sources
cited in each function.  It exists only to *check* the HIP implementation in
``jaccpot_amd``: the only importers allowed are ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``.
Nothing in ``jaccpot_amd`` imports it, and the product path never falls back
to it.

Modules
-------
harmonics  -- packed layout, solid harmonics, P2M/L2P, Dehnen B/Q matrices,
              rotations, z-translations, M2M/M2L/L2L operators.
topology   -- the fixed static-radix topology spec (Morton sort, leaf buckets,
              binary radix tree, COM geometry, dual-tree far/near lists).
fmm        -- the full fixed-order FMM (P2M -> M2M -> M2L -> L2L -> L2P + P2P).
direct     -- the float64 O(N^2) direct-sum oracle.
cfmm       -- ctypes wrapper of cfmm.cpp, a C++/OpenMP restatement of the
              fmm numerics (float64 or float32, optional per-pair orders) on a
              given topology: the oracle at large N and the CPU baseline.
labels     -- per-pair truncation-error labels of the learned order.
mlp        -- the order network's forward/backward in NumPy.
ics        -- seeded initial-condition generators for the benchmark configs.

Parity status of each function is recorded in its docstring ("pinned by ...").
"""
