"""MI355X-native IPR hot path (OpenLPT): HIP kernels behind a C-ABI."""
__version__ = "0.1.0"
