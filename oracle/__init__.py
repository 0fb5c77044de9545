"""CPU parity oracle of the Sailfish split-explicit hot path.

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by the
cpu_baseline leg of bench.py, always as the checker.  The product package
(sailfish_amd) never imports it.
"""
