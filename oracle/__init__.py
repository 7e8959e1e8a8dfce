"""FP64 CPU oracle of the Object-MD force path: test infrastructure only
(tests/, the smoke check and bench.py's CPU-baseline leg)."""
