"""CPU oracle for the SLEEK codec -- test infrastructure only.

Only the tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this package.  The product (``sleek_amd``) never does.
"""
