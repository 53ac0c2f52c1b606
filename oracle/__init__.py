"""CPU oracle for the HDDM full-wave path (test infrastructure only).

Only tests, the smoke check and the CPU-baseline leg of the benchmark import
this package; the product path never does.
"""
