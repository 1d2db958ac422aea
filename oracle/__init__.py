"""CPU oracle for the ensemble membrane solid step - TEST INFRASTRUCTURE ONLY.

This package restates CVFES's CPU solid path (``optimizedSolidAssemble.pyx``, ``sparse.py``,
``physicsSolver.py``) so that the MI355X implementation in ``cvfes_amd`` can be checked against it.
Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import
or run anything here, and only as the checker / the reported CPU baseline - never as part of the
product path.

Pinning status of every function is recorded in ``solid_oracle.py``'s module docstring.
"""
