"""MI355X-native ISLBM periodic-hill solver: gfx950 kernels behind a C ABI (include/islbm.h).

Import with importlib (the folder name is not a Python identifier):

    pkg = importlib.import_module("lattie-boltzmann-method-note_of_paperreview_amd")
    with pkg.Solver(pkg.default_params(precision="fp32")) as s:
        s.init_equilibrium(u0=(0.0, 0.05, 0.0))
        s.step(100)
"""
from .islbm import (  # noqa: F401
    BF16, FP16, FP32, FP64, KERNEL_AUTO, KERNEL_PERNODE, KERNEL_TILED, KERNEL_TILED_WIDE, KERNEL_TWOPASS, MESH_BISECT,
    MESH_STRETCH,
    REQ_ACTIVE, REQ_DONE, REQ_FAILED, REQ_QUEUED, STAT_NAMES, Batch, ISLBMError, Params, Solver, copy_bandwidth, default_params,
    lib, rel_l2, rel_l2_vec, required_bytes,
)
