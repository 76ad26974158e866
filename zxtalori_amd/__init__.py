"""QVM orientation relaxation and spacing on SO(3)/K for the AMD MI355X (gfx950).

The Python surface of the reference's zxtalori package (Grid, haar, from_quaternions,
load, generate) over libzxtalori.so: hand-written HIP kernels behind a C ABI.
"""
__version__ = "0.3.0"
from .api import GROUPS, Grid, from_quaternions, generate, haar, hopf_frequency, hopf_points, load, library_path

__all__ = ["GROUPS", "Grid", "generate", "from_quaternions", "load", "haar", "hopf_points", "hopf_frequency",
           "library_path"]
