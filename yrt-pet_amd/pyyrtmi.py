"""The C++ host layer's pybind11 module (`_pyyrtmi`, built in-tree) under the
reference's pyyrtpet names for this path.

torch (when installed) is imported first: it ships its own HIP runtime as
libamdhip64.so, and a process that loads the system's libamdhip64.so.7 (this
library's) before it ends up with two HIP runtimes, of which the one that
initialises second finds no device.  With torch loaded first, the loader
gives this library torch's runtime, and one process holds one runtime.
"""
try:
    import torch  # noqa: F401
except ImportError:  # no torch: the system runtime serves alone
    pass

from _pyyrtmi import *  # noqa: F401,F403
from _pyyrtmi import YrtmiError, __doc__  # noqa: F401
