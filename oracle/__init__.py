"""CPU oracle of the explicit residual/update path — test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this package,
and only as the checker.  The product (mose_amd) never imports it.
"""
