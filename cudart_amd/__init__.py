"""cudart_amd — MI355X (gfx950) mesh-tree build and closest-hit trace.

The product is the C ABI in include/rk.h (cudart_amd/lib/librk.so) and the C++
drop-in classes in include/renderkit/ (cudart_amd/lib/librenderkit.so).
`cudart_amd.native` binds the C ABI with ctypes; `cudart_amd.device` wraps it
for torch device tensors.
"""
__all__ = ["native"]
