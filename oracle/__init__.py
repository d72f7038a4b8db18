"""CPU oracle of the counting path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may import this package, and only as the
checker. The product (et-miner_amd/et_miner) never imports it.
"""
