"""CPU oracle for the exhaustive scan (test infrastructure only).

Imported only by ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py``; never by the ``pytme_amd`` package.
"""
