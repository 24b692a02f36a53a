"""CPU oracle for the IF-RK4 pseudo-spectral step (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
this package.  The product package ``pyspectral_amd`` never does.
"""
