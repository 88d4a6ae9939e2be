"""CPU oracle for the MOIL energy/force path (test infrastructure only).

Nothing here is part of the product.  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg may import it, and only as the checker.
"""
