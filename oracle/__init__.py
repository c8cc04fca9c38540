"""Parity oracle for the cardiax forward-Euler path.

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and the cpu_baseline
leg of bench.py may import this package, and only to check results or as the CPU
reference timing. The product package ``cardiax_amd`` never imports it.
"""
