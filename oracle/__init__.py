"""CPU oracle of the LAPS time step (test infrastructure only).

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
as the reference results; the product package laps_amd never imports it.
"""
