"""FP64 CPU oracle for the generalized Lomb-Scargle periodogram via the adjoint NFFT.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package (``cuvarbase_amd``)
imports, calls or links anything under ``oracle/``.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use it,
and only as the checker (or, for ``cpu_baseline``, as the timed CPU port).

Modules
-------
``oracle.gls``   data preparation, frequency grids, FP64 direct sums, GLS power
                 in every mode of the reference (closed form and the literal
                 tau-rotated form), multi-harmonic GLS, least-squares definition,
                 Baluev false-alarm probability.
``oracle.nfft``  FP64 adjoint NFFT by Gaussian gridding (centred band and the
                 reference's one-sided band), brute-force gridding, and the
                 NFFT-based GLS used as the CPU baseline.

Pinning status of every function is listed in the header of each module.
"""
