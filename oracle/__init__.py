"""TEST INFRASTRUCTURE ONLY: CPU restatement of the CUBE P3M step (the oracle).

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
checker; never by the product path.
"""
