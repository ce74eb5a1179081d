"""CPU oracle (test infrastructure only): importable by tests/, smoke() and bench.py's cpu_baseline leg."""
