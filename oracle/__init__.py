"""CPU oracle for the MPDO noisy-circuit update path (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
this package, and only as the checker. The product (the HIP library behind ``include/mpdo.h`` and
its Python host) never imports, links or executes anything under ``oracle/``.
"""
