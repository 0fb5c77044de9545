"""CPU parity oracle of the Sailfish split-explicit hot path.

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by the
cpu_baseline leg of bench.py, always as the reference the GPU results are compared
with (or, for the baseline, the CPU path that is timed).  The product package
(sailfish_amd) never imports it.
"""
