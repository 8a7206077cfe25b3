"""fastembedr_amd -- MI355X (gfx950) backend for the fastEmbedR embedding
optimizers: the host side above the C-ABI of ``libfastembedr_hip.so``.

Public surface (mirrors the reference's ``backend = "hip"`` entry points):
``tsne_knn`` (t-SNE from a KNN graph), ``gradient_diagnostic`` (forces from a
fixed layout), ``TsnePlan`` (resident jobs), ``available``.
"""
from ._lib import HipError, available  # noqa: F401
from .tsne import (Embedding, TsnePlan, auto_parameters, force_diagnostic,  # noqa: F401
                   gradient_diagnostic, grid_size, tsne_knn)

__version__ = "0.1.0"
