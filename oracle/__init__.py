"""TEST ORACLE package: plain-C restatement of the reference's per-stage kernels (esdg_oracle.c).

Test infrastructure only.  Imported by tests/, __graft_entry__.smoke() and bench.py's CPU
baseline, always as the checker or the timed CPU reference, never as the product path.
"""
