"""MI355X-native SIMP inner loop: matrix-free hex8 operator, Galerkin GMG-PCG,
compliance / sensitivity, density filter and OC / MMA updates in hand-written
HIP (libdfmi.so) behind a ctypes C ABI.  See include/dfmi.h."""
__all__ = ["element", "problems", "multigrid", "simp"]
