"""binarytpm_amd: the Thermal stage of the binary-asteroid ATPM on MI355X.

libbtpm.so (hand-written gfx950 HIP kernels + C++ runtime behind the C ABI of
include/btpm.h) does the work; this package holds its ctypes front end
(`thermal`), input generators (`inputs`) and file formats (`formats`).
"""
__all__ = ["thermal", "inputs", "formats"]
