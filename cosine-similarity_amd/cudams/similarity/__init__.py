"""Similarity plug-ins (reference cudams/similarity/__init__.py)."""
from .CPUParallelCosineGreedy import CPUParallelCosineGreedy
from .CudaCosineGreedy import CudaCosineGreedy
from .CudaFingerprintSimilarity import CudaFingerprintSimilarity

__all__ = ["CudaCosineGreedy", "CudaFingerprintSimilarity", "CPUParallelCosineGreedy"]
