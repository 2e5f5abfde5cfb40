"""ORACLE -- TEST INFRASTRUCTURE ONLY.

CPU restatements of fairSIM's 2D SIM reconstruction path used as the parity
checker: ``java_order.c`` (FP32, Java operation order, built into
liboracle.so by the Makefile here) and ``fp64.py`` (FP64 NumPy).  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package; the product (fairsim-vigor_amd/) never does.
"""
