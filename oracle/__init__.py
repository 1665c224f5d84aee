"""CPU oracle of the NDTensors contraction path -- test infrastructure only.

Imported by the test-suite, the smoke check and the CPU-baseline leg of the
benchmark, never by the product (the `itg` package / libitg.so).
"""
