"""CPU oracle of the reference's resistive-MHD RK2 step (test infrastructure
only: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
compare against it; never part of the product path)."""
