"""The oracle: CPU restatement of the reference's tiled SpGEMM path.

TEST INFRASTRUCTURE ONLY -- imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker; the product never imports it.
"""
