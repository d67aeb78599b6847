"""Build and load the in-tree libraries: the HIP library (`libsad_hip.so`, gfx950)
and the host-side mesh packer (`libsad_pack.so`).

The library is compiled with hipcc directly (no JIT cache), in-tree, so the same
file is what tests, the smoke check and the benchmark load.  FP32 semantics are
pinned by the flags: no contraction, correctly rounded divide/sqrt, denormals kept,
packed-FP32 instructions disabled.
"""

import ctypes as ct
import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
SOURCE = HERE / "csrc" / "sad_hip.hip"
LIBRARY = Path(os.environ.get("SAND_DEM_HIP_LIBRARY", HERE / "libsad_hip.so"))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def compile_library(output=HERE / "libsad_hip.so", defines=()):
    """Compile csrc/ into `output` (always rebuilds); `defines` are -D tunables."""
    tmp = Path(str(output) + ".tmp")
    cmd = [HIPCC, *FLAGS, *(f"-D{d}" for d in defines), "-o", str(tmp), str(SOURCE)]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode:
        raise RuntimeError(f"hipcc failed ({result.returncode}):\n{result.stderr[-4000:]}")
    tmp.replace(output)
    return output


PACK_SOURCE = HERE / "csrc" / "sad_pack.cpp"
PACK_LIBRARY = HERE / "libsad_pack.so"
INCLUDE = HERE.parent / "include"
CXX = os.environ.get("CXX", "g++")


def compile_packer(output=PACK_LIBRARY):
    """Compile the host-side mesh packer (`csrc/sad_pack.cpp`, plain C++; always
    rebuilds).  No fast-math: bounds are rounded outward with nextafter."""
    tmp = Path(str(output) + ".tmp")
    cmd = [CXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
           f"-I{INCLUDE}", "-o", str(tmp), str(PACK_SOURCE)]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode:
        raise RuntimeError(f"{CXX} failed ({result.returncode}):\n{result.stderr[-4000:]}")
    tmp.replace(output)
    return output


_PACK = None
PACK_EXPORTED = ("sad_pack_node_count", "sad_pack_bvh", "sad_pack_clusters", "sad_sdf_grid")


def packer():
    """The loaded mesh-packer library; raises if it has not been built."""
    global _PACK
    if _PACK is None:
        if not PACK_LIBRARY.exists():
            raise FileNotFoundError(f"{PACK_LIBRARY} is missing: build it with __graft_entry__.build() first")
        lib = ct.CDLL(str(PACK_LIBRARY))
        u32, vp = ct.c_uint32, ct.c_void_p
        lib.sad_pack_node_count.restype, lib.sad_pack_node_count.argtypes = u32, [u32, u32]
        lib.sad_pack_bvh.restype = ct.c_int
        lib.sad_pack_bvh.argtypes = [vp, vp, u32, u32, u32, vp, vp, vp, vp, ct.POINTER(u32)]
        lib.sad_pack_clusters.restype = ct.c_int
        lib.sad_pack_clusters.argtypes = [vp, u32, u32, u32, vp, vp, vp, vp, vp]
        lib.sad_sdf_grid.restype = ct.c_int
        lib.sad_sdf_grid.argtypes = [vp, vp, u32, vp, ct.c_double, vp, u32, vp]
        _PACK = lib
    return _PACK


_LIB = None


def library():
    """The loaded HIP library; raises if it has not been built (no fallback)."""
    global _LIB
    if _LIB is None:
        if not LIBRARY.exists():
            raise FileNotFoundError(f"{LIBRARY} is missing: build it with __graft_entry__.build() first")
        lib = ct.CDLL(str(LIBRARY))
        vp, u32, u64p, dp = ct.c_void_p, ct.c_uint32, ct.POINTER(ct.c_uint64), ct.POINTER(ct.c_double)
        vpp = ct.POINTER(ct.c_void_p)
        signatures = {
            "sad_hip_scratch_create": (vp, []),
            "sad_hip_scratch_destroy": (None, [vp]),
            "sad_hip_scratch_stats": (ct.c_int, [vp, u64p]),
            "sad_hip_dispatch": (ct.c_int, [ct.c_char_p, u32, vpp, u64p, dp, vp, vp]),
            "sad_hip_advance": (ct.c_int, [u32, u32, vpp, u64p, dp, vpp, u64p, vp, vp, ct.c_int]),
            "sad_hip_advance_static": (ct.c_int, [u32, u32, vpp, u64p, dp, vpp, u64p, vp]),
            "sad_hip_probe": (ct.c_int, [ct.c_char_p, ct.POINTER(u32), vpp, u64p, dp, vp, vp, vp]),
            "sad_hip_profile": (ct.c_int, [vp, ct.c_int]),
            "sad_hip_profile_read": (ct.c_int, [vp, dp, u64p, u32]),
            "sad_hip_kernel_name": (ct.c_char_p, [u32]),
            "sad_hip_scan_kernel": (ct.c_int, [ct.c_char_p, u32, vp, vp, vp, vp]),
            "sad_hip_smoke": (ct.c_int, [u32, vp, vp, vp]),
            "sad_hip_normal_probe": (ct.c_int, [u32, ct.c_float, ct.c_float, ct.c_float, vp, vp, vp]),
            "sad_hip_reduce_kernel": (ct.c_int, [ct.c_char_p, u32, u32, u32, vp, ct.c_uint64, vp, ct.c_uint64, vp]),
            "sad_hip_tune": (ct.c_int, [vp, ct.c_char_p, ct.c_int64, ct.POINTER(ct.c_int64)]),
            "sad_hip_invalidate": (ct.c_int, [vp]),
            "sad_hip_nn_configure": (ct.c_int, [vp, ct.c_int, ct.c_int, vp, vp, u32]),
            "sad_hip_nn_grad": (ct.c_int, [vp, vp, vp, vp, u32, u32, vp, vp, vp, vp, ct.c_float, vp]),
            "sad_hip_nn_adam": (ct.c_int, [vp, vp, vp, vp, vp, u32, u32, ct.c_float, ct.c_float, ct.c_float,
                                           ct.c_float, u32, vp]),
            "sad_hip_nn_pack": (ct.c_int, [vp, vp, vp, u32, vp]),
            "sad_hip_nn_infer": (ct.c_int, [vp, vp, u32, vp, vp, vp, vp]),
            "sad_hip_nn_sample": (ct.c_int, [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
            "sad_hip_mfma_probe": (ct.c_int, [vp, vp, vp, ct.c_int, vp]),
            "sad_hip_last_error": (ct.c_char_p, []),
            "sad_hip_abi_version": (ct.c_int, []),
        }
        for name, (res, args) in signatures.items():
            fn = getattr(lib, name)
            fn.restype, fn.argtypes = res, args
        _LIB = lib
    return _LIB


EXPORTED = (
    "sad_hip_scratch_create",
    "sad_hip_scratch_destroy",
    "sad_hip_scratch_stats",
    "sad_hip_dispatch",
    "sad_hip_advance",
    "sad_hip_advance_static",
    "sad_hip_probe",
    "sad_hip_profile",
    "sad_hip_profile_read",
    "sad_hip_kernel_name",
    "sad_hip_scan_kernel",
    "sad_hip_reduce_kernel",
    "sad_hip_smoke",
    "sad_hip_normal_probe",
    "sad_hip_tune",
    "sad_hip_invalidate",
    "sad_hip_nn_configure",
    "sad_hip_nn_grad",
    "sad_hip_nn_adam",
    "sad_hip_nn_pack",
    "sad_hip_nn_infer",
    "sad_hip_nn_sample",
    "sad_hip_mfma_probe",
    "sad_hip_last_error",
    "sad_hip_abi_version",
)
