"""The C-ABI libraries load and export every entry point the header declares
(no compute call: this runs without a GPU)."""
import os
import re
import subprocess

import pytest

from mahyco_amd import GPU_LIB, abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mahyco_gpu.h")


def declared(prefix="mhg_"):
    src = open(HEADER).read()
    return sorted(set(re.findall(r"\b(" + prefix + r"[a-z_0-9]+)\s*\(", src)))


def exported(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True,
                         check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if " T " in ln}


def test_gpu_library_exports_header():
    assert os.path.exists(GPU_LIB), "build first: python -c 'import __graft_entry__ as g; g.build()'"
    names = declared()
    assert len(names) >= 30
    missing = [n for n in names if n not in exported(GPU_LIB)]
    assert not missing, missing


def test_gpu_library_loads_and_binds():
    lib = abi.Library(GPU_LIB, "mhg_")
    assert lib.abi_version() == 1
    o = lib.options()
    assert o.cfl == 0.3 and o.ordre_projection == 2 and o.nb_bc == 6


def test_oracle_exports_same_interface():
    import helpers as H
    names = [n.replace("mhg_", "mho_") for n in declared()]
    have = exported(H.ORACLE_LIB)
    gpu_only = ("mho_run_async", "mho_synchronize", "mho_kernel_times", "mho_autotune",
                "mho_variant_count", "mho_select_variant")
    # the batch form, the network and the per-step inference exist on the GPU only
    gpu_only += ("mho_learned_alpha", "mho_learned_features", "mho_set_limiter_mlp",
                 "mho_set_learned_interface_alpha")
    missing = [n for n in names if n not in have and n not in gpu_only and not n.startswith("mho_batch_")
               and not n.startswith("mho_mlp_")]
    assert not missing, missing


def test_options_struct_matches_header():
    import ctypes as C
    # sizeof(mhg_options) as compiled vs ctypes mirror
    lib = abi.Library(GPU_LIB, "mhg_")
    o = abi.Options()
    lib.default_options(C.byref(o))
    assert o.remap_threshold == 1e-10 and o.is_euler_scheme == 1 and o.visc_quad_coeff == -1.0


def test_batch_create_rejects_mixed_problems():
    """mhg_batch_create validates every problem before any device call."""
    import ctypes as C
    lib = abi.Library(GPU_LIB, "mhg_")
    create = lib.optional["batch_create"]

    def make(n, dims=(2, 2), envs=(1, 1), prec=(0, 0)):
        meshes = (abi.MeshDesc * 2)(*[abi.mesh_desc(d, [8, 4, 4], [0, 0, 0], [1., .5, .5]) for d in dims])
        opts = (abi.Options * 2)()
        for i in range(2):
            opts[i] = lib.options(precision=prec[i])
            abi.set_bcs(opts[i], abi.reflecting_walls(dims[i]))
            abi.set_envs(opts[i], [(abi.EOS_PERFECT_GAS, 1.4, 0.)] * envs[i])
        h = C.c_void_p()
        return create(n, meshes, opts, C.byref(h))

    assert make(0) == abi.E_ARG
    assert make(2, dims=(2, 3)) == abi.E_UNSUPPORTED
    assert make(2, envs=(1, 2)) == abi.E_UNSUPPORTED
    assert make(2, envs=(4, 4)) == abi.E_UNSUPPORTED
    assert make(2, prec=(0, 1)) == abi.E_UNSUPPORTED


def test_create_rejects_meshes_beyond_31_bit_indices():
    """mhg_create refuses a mesh with 2^31 nodes or more before any device
    call (the kernels split cell and node indices in 32-bit arithmetic); a
    mesh just below the limit passes that check (it is not created here)."""
    import ctypes as C
    lib = abi.Library(GPU_LIB, "mhg_")
    o = lib.options()
    abi.set_bcs(o, abi.reflecting_walls(2))
    abi.set_envs(o, [(abi.EOS_PERFECT_GAS, 1.4, 0.)])
    h = C.c_void_p()
    big = abi.mesh_desc(2, [46341, 46341], [0, 0], [1., 1.])  # 46342^2 > 2^31 nodes
    assert lib.create(C.byref(big), C.byref(o), C.byref(h)) == abi.E_UNSUPPORTED
    big3 = abi.mesh_desc(3, [1290, 1290, 1290], [0, 0, 0], [1., 1., 1.])  # 1291^3 > 2^31
    o3 = lib.options()
    abi.set_bcs(o3, abi.reflecting_walls(3))
    abi.set_envs(o3, [(abi.EOS_PERFECT_GAS, 1.4, 0.)])
    assert lib.create(C.byref(big3), C.byref(o3), C.byref(h)) == abi.E_UNSUPPORTED
