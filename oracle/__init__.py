"""CPU oracle of the hex8 SIMP solve path -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, as the checker or the timed CPU baseline; the product never does.
"""
