"""ORACLE adapter -- test infrastructure only.

Only tests/, `__graft_entry__.smoke()` and bench.py's `cpu_baseline` leg import this
module, as the checker.  It exposes the C++ CPU restatement (`cpu_backend.cpp`) and,
when built, the reference's own generated CPU collector (`_ref/ref_collector.so`)
through the same duck-typed adapter protocol as the reference CPU backend
(`src/sand_atlas_dem/cpu.py:431-689`): host NumPy buffers, encoders that record
operations, serial execution at submit.
"""

import ctypes as ct
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIBRARY = HERE / "liboracle.so"
LIBRARY64 = HERE / "liboracle64.so"
REF_COLLECTOR = HERE / "_ref" / "ref_collector.so"
# Slots the restatement reads as integers (cpu_backend.cpp `bind`); every other
# slot holds FP32 values, which the FP64 build reads as doubles.
INTEGER_SLOTS = frozenset((
    "control", "shapeId", "sampleInfo", "meshRoots", "meshHullInfo", "meshLinks", "meshFrontier", "meshPairMask",
    "meshClusterInfo", "meshClusterRanges", "meshSampleIndices", "binCounts", "binParticles", "pairCounts",
    "pairOffsets", "pairs", "contactCounts", "contactOffsets", "meta", "oldMeta", "oldHash", "hash", "heads", "next",
    "wallBoxIds", "grids",
))


def build(reference=True):
    """Compile the oracle, its FP64 twin, and the reference collector when the
    reference exists."""
    targets = ["liboracle.so", "liboracle64.so"] + (["ref"] if reference else [])
    subprocess.run(["make", "-s", "-C", str(HERE), *targets], check=True)


def _load(path, symbol, argtypes):
    lib = ct.CDLL(str(path))
    fn = getattr(lib, symbol)
    fn.argtypes, fn.restype = argtypes, ct.c_int
    return lib, fn


class Buffer:
    def __init__(self, array=None, *, count=None, stride=4):
        self.array = (
            np.ascontiguousarray(array).copy().view(np.uint8).reshape(-1)
            if array is not None
            else np.zeros(max(4, count * stride), np.uint8)
        )
        self.size = self.array.nbytes

    def copy_from_numpy(self, array):
        data = np.ascontiguousarray(array).view(np.uint8).reshape(-1)
        if data.size != self.size:
            raise ValueError("CPU buffer upload size mismatch")
        self.array[:] = data

    def to_numpy(self):
        return self.array.copy()


class WideBuffer(Buffer):
    """FP32-layout buffer with an FP64 shadow for the FP64 build.

    The FP32 bytes stay the buffer's face (reads, uploads, integer use); once the
    buffer is bound to an FP32 slot its values live in `wide` (float64, one double
    per float) and the FP32 face is refreshed from it, rounded, when read.
    """

    def __init__(self, array=None, *, count=None, stride=4):
        super().__init__(array, count=count, stride=stride)
        self.wide = None
        self.stale = False

    def widen(self):
        if self.wide is None:
            self.wide = self.array[: self.size // 4 * 4].view(np.float32).astype(np.float64)
        self.stale = True  # a kernel may write it
        return self.wide

    def sync(self):
        if self.wide is not None and self.stale:
            self.array[: self.wide.size * 4] = self.wide.astype(np.float32).view(np.uint8)
            self.stale = False

    def copy_from_numpy(self, array):
        super().copy_from_numpy(array)
        self.wide, self.stale = None, False

    def to_numpy(self):
        self.sync()
        return super().to_numpy()

    def write_prefix(self, data):
        self.sync()
        self.array[: data.size] = data
        if self.wide is not None:
            n = data.size // 4
            self.wide[:n] = self.array[: 4 * n].view(np.float32)

    def read64(self):
        """The values at full precision (FP32 face if never bound as floats)."""
        return self.wide.copy() if self.wide is not None else self.array.view(np.float32).astype(np.float64)


class Encoder:
    def __init__(self):
        self.operations = []

    def copy_buffer(self, dest, dest_offset, source, source_offset, size):
        if min(dest_offset, source_offset, size) < 0 or dest_offset + size > dest.size or source_offset + size > source.size:
            raise ValueError("CPU buffer copy out of bounds")

        def copy():
            wide = [b for b in (dest, source) if getattr(b, "wide", None) is not None]
            if not wide:
                np.copyto(dest.array[dest_offset : dest_offset + size], source.array[source_offset : source_offset + size])
                return
            if (dest_offset | source_offset | size) % 4:
                raise ValueError("FP64 oracle copies whole elements")
            src = source.widen()[source_offset // 4 : (source_offset + size) // 4]
            dest.widen()[dest_offset // 4 : (dest_offset + size) // 4] = src

        self.operations.append(copy)


REDUCTION_ROWS = {"reduceRows": 8, "reduceFabric": 2, "reduceWallRows": 3}


def reduce_groups(name, records, walls, n, dtype):
    """Group totals of evidence rows in the shader's tree order (diagnostic.slang;
    the reference CPU adapter restates it at cpu.py:491-532).

    `records` holds walls x n records of `rows` float4 rows.  Each group of 256
    records (zero padded) is halved pairwise, 128 -> 1: lane i adds lane i + step.
    reduceRows keeps the maximum in row 4 .w; reduceWallRows carries row 0 .w as an
    unsigned count (FP32 bits) summed with wrap-around.
    """
    rows = REDUCTION_ROWS[name]
    groups = (n + 255) // 256
    values = np.zeros((walls, groups * 256, rows, 4), dtype)
    values[:, :n] = records[: walls * n * rows * 4].reshape(walls, n, rows, 4)
    values = values.reshape(walls, groups, 256, rows, 4)
    counts = None
    if name == "reduceWallRows":
        counts = values[..., 0, 3].astype(np.float32).view(np.uint32).copy()
        values[..., 0, 3] = 0
    step = 128
    while step:
        low, high = values[:, :, :step], values[:, :, step : 2 * step]
        maximum = np.maximum(low[..., 4, 3], high[..., 4, 3]) if name == "reduceRows" else None
        low += high
        if maximum is not None:
            low[..., 4, 3] = maximum
        if counts is not None:
            counts[:, :, :step] += counts[:, :, step : 2 * step]
        step //= 2
    totals = values[:, :, 0].copy()
    if counts is not None:
        totals[..., 0, 3] = counts[:, :, 0].view(np.float32)
    return totals


def pointer_table(bindings, order, wide=False):
    if wide:
        arrays = []
        for name in order:
            b = bindings[name]
            arrays.append(b.array if name in INTEGER_SLOTS else b.widen())
        sizes = [bindings[name].size for name in order]  # FP32 byte sizes: element counts are unchanged
    else:
        arrays = [bindings[name].array for name in order]
        sizes = [a.nbytes for a in arrays]
    pointers = (ct.c_void_p * len(arrays))(*(a.ctypes.data for a in arrays))
    return pointers, (ct.c_uint64 * len(arrays))(*sizes)


def make_oracle(package, threads=1, collector="oracle", precision="fp32"):
    """Create the oracle adapter bound to the package's ABI tables.

    collector="reference" routes collectMeshFrontiers through the reference's own
    generated collector (`_ref/ref_collector.so`, cpu_dispatch ABI), everything else
    through the restatement.  precision="fp64" runs the FP64 build on FP64 shadows
    of the FP32 buffers (`read64` gives the full-precision values).
    """
    order = package.engine.BUFFER_ORDER
    config_values = package.engine.config_values
    wide = precision == "fp64"
    if precision not in ("fp32", "fp64") or (wide and collector != "oracle"):
        raise ValueError("precision is fp32 or fp64 (fp64 with the restatement's collector)")
    path = LIBRARY64 if wide else LIBRARY
    if not path.exists():
        build(reference=False)
    lib, dispatch_fn = _load(
        path, "oracle_dispatch",
        [ct.c_char_p, ct.c_uint32, ct.POINTER(ct.c_void_p), ct.POINTER(ct.c_uint64), ct.POINTER(ct.c_double)],
    )
    probe_fn = lib.oracle_probe
    probe_fn.argtypes = [
        ct.c_char_p, ct.POINTER(ct.c_uint32), ct.POINTER(ct.c_void_p), ct.POINTER(ct.c_uint64),
        ct.POINTER(ct.c_double), ct.c_void_p, ct.c_void_p,
    ]
    probe_fn.restype = ct.c_int
    ref_fn = None
    if collector == "reference":
        if not REF_COLLECTOR.exists():
            raise FileNotFoundError("reference collector not built (oracle/_ref)")
        ref_lib, ref_fn = _load(
            REF_COLLECTOR, "cpu_dispatch",
            [ct.c_uint32, ct.POINTER(ct.c_void_p), ct.POINTER(ct.c_uint64), ct.POINTER(ct.c_double), ct.c_void_p],
        )
        create = ref_lib.cpu_scratch_create
        create.restype = ct.c_void_p
        ref_scratch = create()

    buffer_class = WideBuffer if wide else Buffer

    class Oracle:
        name = "cpu"
        buffer_type = buffer_class
        precision = "fp64" if wide else "fp32"

        def __init__(self):
            self.threads = threads
            self.device = self
            self.calls = 0

        def buffer(self, array=None, *, count=None, stride=4):
            return buffer_class(array, count=count, stride=stride)

        def create_command_encoder(self):
            return Encoder()

        def dispatch(self, name, n, encoder=None, *, r):
            def execute():
                pointers, sizes = pointer_table(r, order, wide)
                values = config_values(r["c"], self.threads)
                vals = values.ctypes.data_as(ct.POINTER(ct.c_double))
                if name == "collectMeshFrontiers" and ref_fn is not None:
                    control = r["control"].array.view(np.uint32)
                    count = min(int(control[1]), r["c"]["pairCapacity"])
                    status = ref_fn(count, pointers, sizes, vals, ref_scratch)
                else:
                    status = dispatch_fn(name.encode(), int(n), pointers, sizes, vals)
                self.calls += 1
                if status:
                    raise RuntimeError(f"oracle dispatch {name} failed with status {status}")

            if encoder is None:
                execute()
            else:
                encoder.operations.append(execute)

        def kernel(self, name, module="solver.slang"):
            if module != "diagnostic.slang" or name not in REDUCTION_ROWS:
                raise NotImplementedError(f"oracle kernel {module}:{name}")
            oracle = self

            class Reduction:
                def dispatch(self, *, thread_count, command_encoder=None, **kwargs):
                    n = int(kwargs["N"])
                    walls = int(thread_count[1]) if name == "reduceWallRows" else 1
                    groups = (n + 255) // 256
                    output_count = int(kwargs.get("outputCount", groups))
                    src, dst = kwargs["input"], kwargs["output"]

                    def execute():
                        rows = REDUCTION_ROWS[name]
                        if wide:
                            records, out, dtype = src.widen(), dst.widen(), np.float64
                        else:
                            records, out, dtype = src.array.view(np.float32), dst.array.view(np.float32), np.float32
                        if n == 0:
                            return
                        if records.size < walls * n * rows * 4 or out.size < walls * output_count * rows * 4:
                            raise ValueError(f"{name}: buffers too small")
                        totals = reduce_groups(name, records, walls, n, dtype)
                        target = out[: walls * output_count * rows * 4].reshape(walls, output_count, rows, 4)
                        target[:, :groups] = totals
                        oracle.calls += 1

                    if command_encoder is None:
                        execute()
                    else:
                        command_encoder.operations.append(execute)

            return Reduction()

        def probe(self, name, uniforms, bindings, data, out):
            pointers, sizes = pointer_table(bindings, order)
            values = config_values(bindings["c"], self.threads)
            u = (ct.c_uint32 * len(uniforms))(*uniforms)
            status = probe_fn(
                name.encode(), u, pointers, sizes, values.ctypes.data_as(ct.POINTER(ct.c_double)),
                None if data is None else data.ctypes.data, out.ctypes.data,
            )
            if status:
                raise RuntimeError(f"oracle probe {name} failed with status {status}")
            return out

        def submit(self, encoder):
            for operation in encoder.operations:
                operation()
            encoder.operations.clear()

        def wait(self):
            pass

        def read(self, buffer, dtype=np.float32, shape=None):
            if wide:
                buffer.sync()
            array = buffer.array.view(dtype).copy()
            return array if shape is None else array.reshape(shape)

        def read_prefix(self, buffer, count, dtype=np.float32, shape=None):
            if type(count) is not int or count < 0 or count * np.dtype(dtype).itemsize > buffer.size:
                raise ValueError("CPU buffer prefix out of bounds")
            if wide:
                buffer.sync()
            array = buffer.array[: count * np.dtype(dtype).itemsize].view(dtype).copy()
            return array if shape is None else array.reshape(shape)

        def write_prefix(self, buffer, array):
            data = np.ascontiguousarray(array).view(np.uint8).reshape(-1)
            if data.size > buffer.size:
                raise ValueError("CPU buffer prefix out of bounds")
            if wide:
                buffer.write_prefix(data)
            else:
                buffer.array[: data.size] = data

    return Oracle()
