"""necpp_amd -- MI355X-native method-of-moments core with nec2++'s C API.

The product is the shared library ``libnecpp_amd.so`` (host C++ plus HIP
kernels for gfx950) behind include/libnecpp.h and include/necpp_gpu.h.  This
package holds its sources (``csrc/``), the in-tree build (``build.py``) and a
ctypes binding with the reference's function names (``necpp.py``).
"""
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libnecpp_amd.so")


def lib():
    from . import necpp
    return necpp.lib()
