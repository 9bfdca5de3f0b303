"""CPU oracle for the 3D R2C/C2R path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The product
(``gpu_fft_amd`` and ``libgpufft.so``) never calls into it.
"""
