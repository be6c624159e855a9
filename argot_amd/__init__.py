"""MI355X (gfx950) implementation of ARGOT's point-source radiation step.

The product is libargot_hip.so (argot_amd/csrc, C-ABI in include/argot_hip.h);
this package holds its ctypes mirror (abi, _lib), the host-side driver in the
reference's call sequence (radiation) and seeded synthetic inputs (ic).
"""
