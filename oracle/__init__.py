"""CPU oracle of the porerad radiation hot path (test infrastructure only).

Only tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this package, as the checker; the product never does.
"""
