"""CPU oracle for the mode-split step -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker. The product
path (``zhenmode_amd``) never imports it and has no CPU fallback.
"""
