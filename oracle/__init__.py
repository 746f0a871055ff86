"""CPU oracle (test infrastructure only): tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg use it as the checker; the product never does."""
