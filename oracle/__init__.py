"""CPU oracle of chromo's Monte Carlo inner loop -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
"""
