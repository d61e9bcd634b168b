"""The oracle: CPU restatement of kursatyurt/UVLM's FMM path -- test infrastructure only.

Importable by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg; never by
the product package uvlm_amd.
"""
