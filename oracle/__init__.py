"""Test infrastructure only: the CPU oracle for the PNVAP hot path.

Importable only from tests/, __graft_entry__.smoke() and the cpu_baseline leg of
bench.py.  The product package (taurus_vap_amd) never imports it.
"""
