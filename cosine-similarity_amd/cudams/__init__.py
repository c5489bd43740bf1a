"""MI355X (gfx950) spectral similarity with the import path of the reference
package ``cudams`` (tornikeo/cosine-similarity), so existing code keeps its
imports: ``from cudams.similarity import CudaCosineGreedy``.

The compute lives in ``libcsamd.so`` (hand-written HIP, C ABI in
include/cosine_similarity_amd.h).  Build it with
``python cosine-similarity_amd/build.py``.
"""
from .__version__ import __version__

__all__ = ["__version__"]
