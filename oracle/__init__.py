"""Test infrastructure: the CPU oracle of the filament velocity.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
may import this package, and only as the checker or the timed CPU baseline.
The product path (qvort_amd) never imports it.
"""
