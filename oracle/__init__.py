"""CPU oracle of the reference's resistive-MHD RK2 step (test infrastructure
only: used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg as the checker; never part of the product path)."""
