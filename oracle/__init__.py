"""CPU oracle for the MADWAVE3 Chebyshev H.psi path -- TEST INFRASTRUCTURE.

Loaded only by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg, as the checker.  The product (madwave3_amd, libmw3) never imports it.
"""
