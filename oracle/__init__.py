"""CPU oracle for the DiskANN search-and-quantize path (test infrastructure only).

Only tests/, the smoke check and the benchmark's CPU-baseline leg import this package,
always as the checker. The product library never calls it.
"""
