"""CPU oracle for the static Skyrme-HF(+BCS) path -- test infrastructure only.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, as the checker.  The product path never imports this package."""
