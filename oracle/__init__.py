"""CPU oracle of the DG right-hand sides -- test infrastructure only.

Importable only from tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg.  The engine never imports it.
"""
