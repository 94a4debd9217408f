"""Pseudo-spectral operators backed by the HIP engine."""
