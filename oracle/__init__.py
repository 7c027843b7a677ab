"""CPU oracle for the VTI FD path — test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker.  The product (geophysics_amd) never
imports it.
"""
