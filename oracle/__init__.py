"""CPU oracle for the AMRClaw per-level patch update.

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
the cpu_baseline leg of bench.py, always as the checker.  The product package
(amrclaw_amd) never imports it.
"""
