"""Test oracle: the reference's CPU path (oracle/_ref) and a numpy restatement.

TEST INFRASTRUCTURE ONLY -- imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product (chucky_amd/).
"""
