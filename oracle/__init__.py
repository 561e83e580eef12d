"""FP64 CPU oracle of the reference local-inertial step.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product (adaptflood_amd/)
never imports it.
"""
