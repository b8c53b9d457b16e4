"""CPU oracle of Sparkle's `cpu` path-tracer hot path — test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this package, and only as
the checker or the timed CPU baseline; the product (sparkle_amd/) never imports it.
"""
