"""The `hip` backend adapter: device buffers in HBM, kernels through the C ABI.

Implements the duck-typed adapter protocol the reference `Engine` drives
(`src/sand_atlas_dem/backend.py:9-121` for Metal, `cpu.py:540-689` for CPU):
buffer / kernel dispatch / command encoder with copy_buffer / submit / wait / read /
read_prefix / write_prefix / doctor.  Device memory and the stream come from
PyTorch-ROCm; every kernel is the hand-written gfx950 code in `csrc/`, reached via
ctypes.  There is no CPU fallback: a missing library or device raises.
"""

import ctypes as ct
import platform

import numpy as np

from . import native
from .engine import BUFFER_ORDER, config_values

SIDE_ORDER = ("meta", "slip", "normalArea", "hash")
# Buffers whose contents the collection stage's temporal-skip anchors depend on
# (template geometry, node samples, body scales, shape ids).  Poses are not among
# them: the anchors bound the motion of the poses themselves.
GEOMETRY = ("shapeId", "body", "sampleInfo", "sample", "meshRoots", "meshHullInfo", "meshPlanes", "meshMin",
            "meshMax", "meshLinks", "meshTriangles", "meshNormals", "meshFrontier", "meshClusterInfo",
            "meshClusterRanges", "meshClusterMin", "meshClusterMax", "meshSampleIndices", "grids", "gridOrigin",
            "distance")


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("The hip backend needs a ROCm GPU; none is visible")
    return torch


class Buffer:
    """A device allocation (uint8 tensor) with the reference Buffer interface."""

    def __init__(self, backend, array=None, *, count=None, stride=4):
        torch = backend.torch
        self.backend = backend
        if array is not None:
            host = np.ascontiguousarray(array).view(np.uint8).reshape(-1)
            self.tensor = torch.from_numpy(host.copy()).to(backend.torch_device)
        else:
            self.tensor = torch.zeros(max(4, count * stride), dtype=torch.uint8, device=backend.torch_device)
        self.size = int(self.tensor.numel())
        self.ptr = int(self.tensor.data_ptr())
        backend.uploaded(self)  # a new allocation may reuse a freed geometry buffer's address

    def copy_from_numpy(self, array):
        data = np.ascontiguousarray(array).view(np.uint8).reshape(-1)
        if data.size != self.size:
            raise ValueError("HIP buffer upload size mismatch")
        import torch

        self.backend.uploaded(self)
        self.tensor.copy_(torch.from_numpy(data.copy()))

    def to_numpy(self):
        return self.tensor.cpu().numpy().copy()


class Encoder:
    """Records copies and kernel launches; `submit` issues them in order."""

    def __init__(self):
        self.operations = []

    def copy_buffer(self, dest, dest_offset, source, source_offset, size):
        if min(dest_offset, source_offset, size) < 0 or dest_offset + size > dest.size or source_offset + size > source.size:
            raise ValueError("HIP buffer copy out of bounds")
        self.operations.append(
            lambda: dest.tensor[dest_offset : dest_offset + size].copy_(source.tensor[source_offset : source_offset + size])
        )


def _table(bindings):
    ptrs = (ct.c_void_p * len(BUFFER_ORDER))(*(bindings[name].ptr for name in BUFFER_ORDER))
    sizes = (ct.c_uint64 * len(BUFFER_ORDER))(*(bindings[name].size for name in BUFFER_ORDER))
    return ptrs, sizes


class Kernel:
    """`m.kernel(name, module)` of the adapter protocol (gpu.py:75-90 Scan, 529-530).

    solver.slang kernels go through the library's dispatch table; scan.slang's
    blockScan/addBlocks through `sad_hip_scan_kernel`; diagnostic.slang's
    reduceRows/reduceFabric/reduceWallRows through `sad_hip_reduce_kernel`;
    normal.slang's normalProbe through `sad_hip_normal_probe`.
    """

    MODULES = {
        "solver.slang": None,
        "scan.slang": {"blockScan", "addBlocks"},
        "diagnostic.slang": {"reduceRows", "reduceFabric", "reduceWallRows"},
        "normal.slang": {"normalProbe"},
    }

    def __init__(self, backend, name, module):
        if module not in self.MODULES:
            raise NotImplementedError(f"{module} kernels are not provided by the hip backend")
        if self.MODULES[module] is not None and name not in self.MODULES[module]:
            raise NotImplementedError(f"{module} has no kernel {name}")
        self.m, self.name, self.module = backend, name, module

    def dispatch(self, *, thread_count, command_encoder=None, vars=None, **kwargs):
        if self.module == "solver.slang":
            r = (vars or {}).get("r", kwargs.get("r"))
            return self.m.dispatch(self.name, int(thread_count[0]), command_encoder, r=r)
        n = int(kwargs["N"])
        if self.module == "normal.slang":
            src, dst = kwargs["input"], kwargs["output"]
            if min(src.size, dst.size) < 16 * n:
                raise ValueError("normalProbe buffers hold fewer than N rows")
            args = (float(kwargs["stiffness"]), float(kwargs["viscosity"]), float(kwargs["weight"]))

            def probe():
                self.m._check(self.m.lib.sad_hip_normal_probe(n, *args, ct.c_void_p(src.ptr), ct.c_void_p(dst.ptr),
                                                              self.m.stream), "normalProbe")

            return probe() if command_encoder is None else command_encoder.operations.append(probe)
        if self.module == "diagnostic.slang":
            return self._reduce(n, thread_count, command_encoder, kwargs)
        if self.name == "blockScan":
            args = (kwargs["input"], kwargs["output"], kwargs["totals"])
        elif self.name == "addBlocks":
            args = (None, kwargs["output"], kwargs["offsets"])
        else:
            raise NotImplementedError(f"scan kernel {self.name}")

        def execute():
            src, dst, aux = args
            self.m._check(self.m.lib.sad_hip_scan_kernel(
                self.name.encode(), n, None if src is None else ct.c_void_p(src.ptr), ct.c_void_p(dst.ptr),
                ct.c_void_p(aux.ptr), self.m.stream), self.name)

        if command_encoder is None:
            execute()
        else:
            command_encoder.operations.append(execute)

    def _reduce(self, n, thread_count, command_encoder, kwargs):
        walls = int(thread_count[1]) if self.name == "reduceWallRows" else 1
        groups = (n + 255) // 256
        output_count = int(kwargs.get("outputCount", groups))
        if int(thread_count[0]) != 256 * groups:
            raise ValueError(f"{self.name}: {thread_count[0]} threads for {n} records")
        src, dst = kwargs["input"], kwargs["output"]

        def execute():
            self.m._check(self.m.lib.sad_hip_reduce_kernel(
                self.name.encode(), n, walls, output_count, ct.c_void_p(src.ptr), src.size, ct.c_void_p(dst.ptr),
                dst.size, self.m.stream), self.name)

        if command_encoder is None:
            execute()
        else:
            command_encoder.operations.append(execute)


class HipBackend:
    name = "hip"

    def __init__(self, device=0, graph=False, temporal_skip=True):
        self.torch = _torch()
        self.torch_device = self.torch.device("cuda", device)
        self.torch.cuda.set_device(self.torch_device)
        self.device = self  # the Engine reaches encoders and wait() via m.device (gpu.py:581, 727)
        self.lib = native.library()
        self.scratch = self.lib.sad_hip_scratch_create()
        if not self.scratch:
            raise MemoryError("HIP scratch creation failed")
        self.graph = graph
        self.buffer_type = Buffer
        self.kernels = {}
        # Temporal node skip: identical results, fewer exact queries between pose
        # changes.  A new buffer, or an upload into a buffer that a launch then binds
        # as geometry, drops the anchors, so stored distances never outlive the
        # geometry they were measured on.
        self.dirty = set()
        self.tune("temporal_skip", int(bool(temporal_skip)))

    def invalidate(self):
        """Forget the collection stage's temporal-skip anchors."""
        self._check(self.lib.sad_hip_invalidate(self.scratch), "invalidate")

    def uploaded(self, buffer):
        """Note a host upload; it drops the anchors if a launch binds it as geometry."""
        self.dirty.add(buffer.ptr)

    def _settle(self, bindings):
        if self.dirty:
            if any(name in bindings and bindings[name].ptr in self.dirty for name in GEOMETRY):
                self.invalidate()
            self.dirty.clear()

    def __del__(self):
        lib, scratch = getattr(self, "lib", None), getattr(self, "scratch", None)
        if lib is not None and scratch:
            try:
                self.torch.cuda.synchronize()
            except Exception:
                pass
            lib.sad_hip_scratch_destroy(scratch)
            self.scratch = None

    # ---- adapter protocol
    def buffer(self, array=None, *, count=None, stride=4):
        return Buffer(self, array, count=count, stride=stride)

    def create_command_encoder(self):
        return Encoder()

    def kernel(self, name, module="solver.slang"):
        key = (module, name)
        if key not in self.kernels:
            self.kernels[key] = Kernel(self, name, module)
        return self.kernels[key]

    @property
    def stream(self):
        return ct.c_void_p(self.torch.cuda.current_stream(self.torch_device).cuda_stream)

    def _check(self, status, what):
        if status:
            raise RuntimeError(f"{what} failed ({status}): {self.lib.sad_hip_last_error().decode()}")

    def dispatch(self, name, n, encoder=None, *, r=None, **buffers):
        if name == "smoke":  # its own bindings (solver.slang:4-10; backend.py:106)
            count = int(buffers.get("N", n))
            a, b = buffers["a"], buffers["b"]
            if min(a.size, b.size) < 4 * count:
                raise ValueError("smoke buffers hold fewer than N floats")

            def smoke():
                self._check(self.lib.sad_hip_smoke(count, ct.c_void_p(a.ptr), ct.c_void_p(b.ptr), self.stream),
                            "dispatch smoke")

            return smoke() if encoder is None else encoder.operations.append(smoke)
        if r is None:
            raise TypeError("solver kernels take the Resources bindings as r=")

        def execute():
            self._settle(r)
            ptrs, sizes = _table(r)
            values = config_values(r["c"])
            self._check(
                self.lib.sad_hip_dispatch(
                    name.encode(), int(n), ptrs, sizes, values.ctypes.data_as(ct.POINTER(ct.c_double)),
                    self.scratch, self.stream,
                ),
                f"dispatch {name}",
            )

        if encoder is None:
            execute()
        else:
            encoder.operations.append(execute)

    def encode_steps(self, encoder, steps, active, bindings, history_buffers):
        """Whole timesteps in one native call (per-step order of gpu.py:714-726)."""

        def execute():
            self._settle(bindings)
            ptrs, sizes = _table(bindings)
            values = config_values(bindings["c"])
            sides = [history_buffers[s][name] for s in (0, 1) for name in SIDE_ORDER]
            sp = (ct.c_void_p * 8)(*(b.ptr for b in sides))
            ss = (ct.c_uint64 * 8)(*(b.size for b in sides))
            self._check(
                self.lib.sad_hip_advance(
                    steps, active, ptrs, sizes, values.ctypes.data_as(ct.POINTER(ct.c_double)), sp, ss,
                    self.scratch, self.stream, int(self.graph),
                ),
                "advance",
            )

        encoder.operations.append(execute)

    def probe(self, name, uniforms, bindings, data, out):
        """Run a query probe on device; `data`/`out` are host arrays."""
        torch = self.torch
        src = None if data is None else torch.from_numpy(np.ascontiguousarray(data).copy()).to(self.torch_device)
        dst = torch.zeros(out.nbytes, dtype=torch.uint8, device=self.torch_device)
        self._settle(bindings)
        ptrs, sizes = _table(bindings)
        values = config_values(bindings["c"])
        u = (ct.c_uint32 * 4)(*(list(uniforms) + [0] * (4 - len(uniforms))))
        self._check(
            self.lib.sad_hip_probe(
                name.encode(), u, ptrs, sizes, values.ctypes.data_as(ct.POINTER(ct.c_double)),
                None if src is None else ct.c_void_p(src.data_ptr()), ct.c_void_p(dst.data_ptr()), self.stream,
            ),
            f"probe {name}",
        )
        out.view(np.uint8).reshape(-1)[:] = dst.cpu().numpy()
        return out

    def probe_device(self, name, uniforms, bindings, src, dst):
        """Query probe on device tensors (no host round trip)."""
        self._settle(bindings)
        ptrs, sizes = _table(bindings)
        values = config_values(bindings["c"])
        u = (ct.c_uint32 * 4)(*(list(uniforms) + [0] * (4 - len(uniforms))))
        self._check(
            self.lib.sad_hip_probe(
                name.encode(), u, ptrs, sizes, values.ctypes.data_as(ct.POINTER(ct.c_double)),
                None if src is None else ct.c_void_p(src.data_ptr()), ct.c_void_p(dst.data_ptr()), self.stream,
            ),
            f"probe {name}",
        )
        return dst

    def tune(self, name, value=-1):
        """Read (value < 0) or set a collection launch tunable; returns the previous value."""
        prev = ct.c_int64(0)
        self._check(self.lib.sad_hip_tune(self.scratch, name.encode(), int(value), ct.byref(prev)), f"tune {name}")
        return int(prev.value)

    def configure_cull(self, enabled, audit=False, weights=None, blocks=None, templates=0):
        """Switch the learned cull of the collection stage on or off (device tensors)."""
        self._check(
            self.lib.sad_hip_nn_configure(
                self.scratch, int(enabled), int(audit),
                ct.c_void_p(weights.data_ptr()) if enabled else None,
                ct.c_void_p(blocks.data_ptr()) if enabled else None, int(templates),
            ),
            "nn_configure",
        )

    def submit(self, encoder):
        for operation in encoder.operations:
            operation()
        encoder.operations.clear()

    def wait(self):
        self.torch.cuda.synchronize(self.torch_device)

    def read(self, buffer, dtype=np.float32, shape=None):
        array = buffer.tensor.cpu().numpy().view(dtype).copy()
        return array if shape is None else array.reshape(shape)

    def read_prefix(self, buffer, count, dtype=np.float32, shape=None):
        if type(count) is not int or count < 0 or count * np.dtype(dtype).itemsize > buffer.size:
            raise ValueError("HIP buffer prefix out of bounds")
        n = count * np.dtype(dtype).itemsize
        array = buffer.tensor[:n].cpu().numpy().view(dtype).copy()
        return array if shape is None else array.reshape(shape)

    def write_prefix(self, buffer, array):
        data = np.ascontiguousarray(array).view(np.uint8).reshape(-1)
        if data.size > buffer.size:
            raise ValueError("HIP buffer prefix out of bounds")
        if data.size:
            self.uploaded(buffer)
            buffer.tensor[: data.size].copy_(self.torch.from_numpy(data.copy()))

    def stats(self):
        """Cumulative device work counters of this backend's workspace."""
        out = (ct.c_uint64 * 12)()
        self._check(self.lib.sad_hip_scratch_stats(self.scratch, out), "stats")
        keys = ("collections", "directed_pairs", "queried_nodes", "contacts", "scratch_bytes", "nn_culled",
                "nn_audited", "nn_misses", "skipped_nodes", "unchanged_winners", "anchored_tasks",
                "derived_rebuilds")
        return dict(zip(keys, (int(v) for v in out)))

    def profile(self, enable=True):
        """Start (clearing) or stop per-kernel HIP-event timing."""
        self._check(self.lib.sad_hip_profile(self.scratch, int(enable)), "profile")

    def profile_read(self):
        """{kernel: (total ms, launches)} accumulated since profiling was enabled."""
        ms = (ct.c_double * 64)()
        launches = (ct.c_uint64 * 64)()
        n = self.lib.sad_hip_profile_read(self.scratch, ms, launches, 64)
        if n < 0:
            raise RuntimeError(self.lib.sad_hip_last_error().decode())
        return {self.lib.sad_hip_kernel_name(k).decode(): (ms[k], int(launches[k])) for k in range(n) if launches[k]}

    def doctor(self):
        """Device report with the compiled-kernel check (backend.py:103-119)."""
        props = self.torch.cuda.get_device_properties(self.torch_device)
        x = np.linspace(-2, 2, 257, dtype=np.float32)
        a, b = self.buffer(x), self.buffer(np.zeros_like(x))
        self.dispatch("smoke", len(x), N=len(x), a=a, b=b)
        result = self.read(b)
        error = float(np.max(np.abs(result - (x * x + 1))))
        if not np.allclose(result, x * x + 1, rtol=1e-6, atol=1e-6):
            raise RuntimeError("HIP compiled kernel smoke check failed")
        return dict(
            format="sand-dem-doctor/1",
            backend="hip",
            fallback=False,
            adapter=props.name,
            gcn_arch=getattr(props, "gcnArchName", ""),
            compute_units=props.multi_processor_count,
            abi_version=int(self.lib.sad_hip_abi_version()),
            numerics="hand-written gfx950 HIP, FP32 compensated state, CPU-order deterministic collection",
            host=platform.processor() or platform.machine(),
            smoke_elements=len(x),
            smoke_passed=True,
            smoke_max_error=error,
        )
