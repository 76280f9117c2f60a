"""ORACLE -- test infrastructure only, never product code.

CPU restatement of the reference's local-DVC kernels (``cpu_ref.c`` + ``cpu_ref.py``) and of
its NCC peak logic (``ncc_ref.py``). Only tests/, the smoke check and the CPU-baseline leg of
the benchmark may import this package.
"""
