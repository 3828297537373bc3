"""CPU oracle for the temporal-motif query path (test infrastructure only).

This package is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
it.  The product path (``everest_amd``) never calls into it and fails loudly
when its HIP library is missing.

Contents
--------
motif.py   restatement of the reference motif reader and pre-decoder
           (src/core/cpu/Graph.cc:198-259, src/core/cpu/data.cc:411-449)
prep.py    numpy restatement of the edge-list loader and CSR build
           (src/core/cpu/data.cc:175-253, src/core/gpu/data.cu:124-348)
brute.py   brute-force count/enumeration from the match definition (tiny inputs)
tm_oracle.c  C restatement of the reference per-root DFS
           (plugins/GPUWorkerU1B.cu:10-263, system/templates/common/*.snp),
           multithreaded, instrumented with the R/X1/X2/P work counters
oracle.py  ctypes wrapper over the compiled tm_oracle.c
"""
