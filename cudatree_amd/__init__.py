"""MI355X-native random forest with CudaTree's API (placeholder during bring-up)."""
__version__ = "0.1.0"
