"""CPU oracle for the lattice-gas engine -- test infrastructure only.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this package.  The engine (``lgca_amd``) never does.
"""
