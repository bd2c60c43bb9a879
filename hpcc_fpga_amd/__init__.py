"""MI355X (gfx950) backend of the HPCC LINPACK hot path.

libhpl_hip.so (built from hpcc_fpga_amd/csrc) implements the C ABI declared in
include/hpl_hip.h; hpcc_fpga_amd.linpack binds it with ctypes, and the C++
host under hpcc_fpga_amd/host is the drop-in LINPACK benchmark executable.
"""
from .linpack import (Batch, Context, HplError, Server, Timing, gefa, gesl,  # noqa: F401
                      load_library)

__all__ = ["Batch", "Context", "HplError", "Server", "Timing", "gefa", "gesl", "load_library"]
