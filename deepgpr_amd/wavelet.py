"""The ``DeepGPR.wavelet`` access path of the reference's quick start
(src/DeepGPR/wavelet.py: ``DeepGPR.wavelet.ricker(...)``) for the source
wavelets this package provides."""
from .preprocessing.wavelets import gaussian, gaussian_derivative, ricker  # noqa: F401

__all__ = ["gaussian", "gaussian_derivative", "ricker"]
