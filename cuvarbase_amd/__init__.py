"""cuvarbase_amd -- generalized Lomb-Scargle periodograms via the adjoint NFFT on
AMD MI355X (gfx950), with the Python API of cuvarbase's GLS path.

Modules mirror the reference: ``lombscargle`` (LombScargleAsyncProcess,
lomb_scargle_async, ...), ``cunfft`` (NFFTAsyncProcess, nfft_adjoint_async),
``core`` (GPUAsyncProcess), ``utils``.  The compute path is the HIP library
``libgls.so`` (C ABI in include/gls.h) loaded with ctypes; no device context is
created at import time.
"""
__version__ = "0.1.0"
