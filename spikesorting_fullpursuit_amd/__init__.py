"""MI355X-native Full Binary Pursuit engine (gfx950 HIP kernels behind a C-ABI)."""
__version__ = "0.1.0"
