"""epcq: HashPoint's point searching and sampling path on AMD MI355X (gfx950).

`epcq.epcq` keeps the Python API of the reference package (CameraSpec,
RenderOptions, NeuralPoints, FeatureQuery, ...) and `epcq.csrc` its op
module, backed by hand-written HIP kernels behind a C ABI.
"""
from .version import __version__  # noqa: F401
