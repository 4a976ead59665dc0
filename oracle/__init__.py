"""ORACLE -- test infrastructure only (CPU restatement of the reference's render
hot path).  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg may import it; the product never does."""
