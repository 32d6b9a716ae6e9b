"""pamtra_amd: PAMTRA's passive microwave forward operator on MI355X.

The pyPamtra API (pyPamtra(), nmlSet, df.addHydrometeor, createProfile,
runPamtra) backed by hand-written HIP kernels for gfx950 behind the C-ABI of
include/pamtra_amd.h.
"""

from . import importer, meteoSI
from .core import pyPamtra
from .descriptorFile import pamDescriptorFile

__all__ = ["pyPamtra", "pamDescriptorFile", "importer", "meteoSI"]
__version__ = "0.1.0"
