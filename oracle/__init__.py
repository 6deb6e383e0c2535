"""CPU oracle of the level-set hot path -- test infrastructure only.

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg;
the product package never imports it.
"""
