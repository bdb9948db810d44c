"""ORACLE — test infrastructure only.

A CPU restatement of the reference's kernels (``refgen``: the text generator,
``runner``: gcc build + ctypes calls).  Imported by ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
only, and there only as the checker; the product never imports it.
"""
