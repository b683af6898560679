"""ORACLE: CPU restatements of the reference's CPU engine, for checking the HIP backend.

Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's CPU baseline import it; the product never does.
"""
