"""CPU oracle for the ADRT transform core -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package.  It is the checker, never the product:
the ``adrt_amd`` package does not import it and has no CPU fallback.

Modules
-------
``oracle.oracle``
    The plain-C restatement (``adrt_oracle.c``) behind NumPy wrappers, plus
    NumPy restatements of the reference's pure-Python routines.
``oracle.ref``
    The reference's own native core compiled from the reference checkout
    (``oracle/_ref/libadrt_ref.so``), used to pin the restatement and as the
    CPU baseline.  Absent when the checkout was not available at build time.
"""
