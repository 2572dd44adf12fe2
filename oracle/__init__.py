"""The oracle: a CPU restatement of the reference engine (md-ref), used only as the checker by
the tests, the smoke check and the benchmark's CPU-baseline leg."""
