"""MI355X-native ensemble shell-membrane solid step (CVFES solid path on gfx950).

Host side: mesh generation (:mod:`.mesh`), block structure (:mod:`.sparse`), GMRF inputs
(:mod:`.gmrf`), the device-resident system (:mod:`.system`) and the CVFES solver-class interface
(:mod:`.physicsSolverSDampGPUs`, :mod:`.physicsSolverGPUs`).  Device side: the HIP kernels behind
the C ABI of ``include/cvfes_mi355x.h`` (``libcvfes_mi355x.so``).
"""
__version__ = "0.1.0"
