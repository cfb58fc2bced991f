"""TEST INFRASTRUCTURE ONLY: the CPU oracle of the tracking path.

Importable by tests/, ``__graft_entry__.smoke()`` and bench.py's
cpu_baseline leg; the product package ``at_amd`` never imports it.
"""
