"""ORACLE — test infrastructure only: a CPU restatement of the reference hal::cpu backend and ring
layer, used by tests/ as the parity checker and by bench.py as the CPU baseline. The product never
imports this package."""
