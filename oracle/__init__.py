"""CPU oracle (test infrastructure): a restatement of stdgpu's OpenMP hash-table engine.

Only tests, the smoke check and the benchmark's CPU-baseline leg import this package.
"""
