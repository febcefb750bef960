"""ORACLE — test infrastructure only (CPU restatement of the reference's path). Imported by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by cudqc_amd."""
