"""CPU oracle of the CURVIB fractional-step core (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package; the HIP product path never does.
"""
