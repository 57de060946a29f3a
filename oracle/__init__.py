"""CPU oracle of the Cyclix CheFSI + AAR hot path -- test infrastructure only
(imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg)."""
