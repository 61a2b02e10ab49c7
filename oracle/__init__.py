"""CPU oracle for the RF demodulation + TBC hot path.

THIS PACKAGE IS TEST INFRASTRUCTURE.  It is an FP64 numpy/scipy restatement of
ld-decode's algorithms (every function cites the reference file:line it
follows) and it is used only as the checker: by ``tests/``, by
``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of ``bench.py``.
The product path (``ld-decode_amd``) never imports, links or executes anything
under ``oracle/``.

Parity status per function (how each one is pinned) is recorded in the module
docstrings of ``oracle.dsp`` and ``oracle.tbc``.
"""
