"""MI355X (gfx950) engine for MEPinterp's k-space hot path.

Wannier interpolation (Fourier GEMM on MFMA), batched Hermitian
eigensolves, velocity rotation and the Berry / Kubo / MEP band sums, behind
the C-ABI of include/mepinterp.h.  Submodules:

* capi     -- ctypes binding of libmepinterp_amd.so (no CPU fallback)
* host     -- core_worker-style driver, fermi_selector, normalisation, NPY
* formats  -- _hr.dat, input.cfg and NPY readers / writers
* models   -- seeded synthetic tight-binding models
* build    -- hipcc build of the library
"""
__all__ = ["capi", "host", "formats", "models", "build"]
