"""Measurement helpers for rocprofv3 counter runs (not used by the library)."""
