"""MI355X-native CommonNN clustering engine.

Drop-in for the ``Clustering.fit(radius_cutoff, cnn_cutoff, member_cutoff)``
path of cnnclustering 0.5.0: the same component recipe (InputData / Metric /
NeighboursGetter / Neighbours / SimilarityChecker / Fitter / Labels) with GPU
components behind a C-ABI (``include/cnn_gpu.h``, ``libcnn_hip.so``).
"""

__version__ = "0.1.0"

from .cluster import Bundle, Clustering  # noqa: E402,F401
from ._types import ClusterParameters, Labels  # noqa: E402,F401
from .report import Record, Summary  # noqa: E402,F401
