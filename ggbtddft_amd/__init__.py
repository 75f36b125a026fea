"""ggbtddft_amd: MI355X-native real-time TDDFT propagator (gfx950 HIP kernels
behind a C ABI, Python host) for atoms in linearly polarised laser fields.

Host modules: grid (Gauss-Lobatto radial DVR), angular (Legendre tables),
atoms, pulse, lda, groundstate (imaginary-time FP64), system (precompute),
engine (the C-ABI propagator), spectrum, config and io (reference-compatible
inputs and outputs).
"""
__version__ = "0.1.0"
