"""CPU oracle for the pair-counting hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke`` and the ``cpu_baseline`` leg of bench.py
may import this package.  The product (``halotools_amd``) never does.
"""
