"""Similarity plug-ins (reference cudams/similarity/__init__.py)."""
from .CudaCosineGreedy import CudaCosineGreedy
from .CudaFingerprintSimilarity import CudaFingerprintSimilarity

__all__ = ["CudaCosineGreedy", "CudaFingerprintSimilarity"]
