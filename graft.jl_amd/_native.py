"""ctypes binding of libgraft_heff.so (the C ABI declared in include/graft_heff.h).

This is the Python counterpart of the ``ccall`` layer a Julia extension would
use.  Loading fails loudly when the library has not been built: there is no
CPU fallback anywhere on the product path.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

MAX_RANK = 12
PREC_FP64, PREC_FP32, PREC_BF16X3 = 0, 1, 2
PRECISIONS = {"fp64": PREC_FP64, "fp32": PREC_FP32, "bf16x3": PREC_BF16X3}

LIB_NAME = "libgraft_heff.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)


class GraftError(RuntimeError):
    """A non-zero status returned by the native library."""

    def __init__(self, status, message):
        super().__init__(f"graft status {status}: {message}")
        self.status = status


class PairStepC(ctypes.Structure):
    _fields_ = [
        ("a", ctypes.c_int32), ("b", ctypes.c_int32), ("dst", ctypes.c_int32),
        ("n_oindA", ctypes.c_int32), ("n_cindA", ctypes.c_int32),
        ("n_oindB", ctypes.c_int32), ("n_cindB", ctypes.c_int32),
        ("oindA", ctypes.c_int32 * MAX_RANK), ("cindA", ctypes.c_int32 * MAX_RANK),
        ("oindB", ctypes.c_int32 * MAX_RANK), ("cindB", ctypes.c_int32 * MAX_RANK),
        ("conjA", ctypes.c_int32), ("conjB", ctypes.c_int32),
        ("n_out_codomain", ctypes.c_int32), ("n_out_domain", ctypes.c_int32),
        ("out", ctypes.c_int32 * MAX_RANK),
    ]


class TensorDescC(ctypes.Structure):
    _fields_ = [
        ("rank", ctypes.c_int32),
        ("dims", ctypes.c_int64 * MAX_RANK),
        ("nsectors", ctypes.c_int32 * MAX_RANK),
        ("sector_dims", ctypes.POINTER(ctypes.c_int64)),
        ("nblocks", ctypes.c_int64),
        ("block_sectors", ctypes.POINTER(ctypes.c_int32)),
        ("block_offsets", ctypes.POINTER(ctypes.c_int64)),
        ("block_strides", ctypes.POINTER(ctypes.c_int64)),
    ]


_lib = None


def lib():
    """Load the native library once; raise if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_NAME} is not built ({LIB_PATH}); run __graft_entry__.build() first. "
            "The effective-Hamiltonian path has no CPU fallback.")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64, u64, dp = (ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64,
                             ctypes.c_uint64, ctypes.POINTER(ctypes.c_double))
    pp = ctypes.POINTER(ctypes.c_void_p)
    sig = {
        "graft_ctx_create": [ctypes.c_int, vp, ctypes.POINTER(vp)],
        "graft_ctx_destroy": [vp],
        "graft_ctx_set_stream": [vp, vp],
        "graft_ctx_set_workspace_cap": [vp, u64],
        "graft_ctx_workspace_bytes": [vp, ctypes.POINTER(u64)],
        "graft_ctx_release_workspace": [vp],
        "graft_synchronize": [vp],
        "graft_malloc": [vp, u64, ctypes.POINTER(vp)],
        "graft_free": [vp, vp],
        "graft_upload": [vp, vp, vp, u64],
        "graft_download": [vp, vp, vp, u64],
        "graft_plan_create": [vp, ctypes.POINTER(PairStepC), i32, i32, i32,
                              ctypes.POINTER(TensorDescC), i32, ctypes.POINTER(TensorDescC),
                              ctypes.POINTER(vp)],
        "graft_plan_create_coupled": [vp, ctypes.POINTER(PairStepC), i32, i32, i32,
                                      ctypes.POINTER(TensorDescC), i32,
                                      ctypes.POINTER(TensorDescC), ctypes.POINTER(i64),
                                      ctypes.POINTER(i32), dp, ctypes.POINTER(vp)],
        "graft_plan_destroy": [vp],
        "graft_plan_info": [vp, ctypes.POINTER(u64), dp, ctypes.POINTER(u64)],
        "graft_plan_device_bytes": [vp, ctypes.POINTER(u64)],
        "graft_plan_step_info": [vp, i32, ctypes.POINTER(i64), ctypes.POINTER(i64), dp,
                                 ctypes.POINTER(i32)],
        "graft_plan_set_tile_config": [vp, i32, i32],
        "graft_plan_set_orientation": [vp, i32, i32],
        "graft_plan_output_blocks": [vp, ctypes.POINTER(i32), ctypes.POINTER(i64),
                                     ctypes.POINTER(i32), ctypes.POINTER(i64)],
        "graft_plan_step_launch": [vp, i32, ctypes.POINTER(i32), ctypes.POINTER(i32),
                                   ctypes.POINTER(i32)],
        "graft_plan_step_elements": [vp, i32, ctypes.POINTER(i64)],
        "graft_plan_step_gemms": [vp, i32, ctypes.POINTER(i64), ctypes.POINTER(i32)],
        "graft_plan_autotune": [vp, vp, pp, vp, i32, i32],
        "graft_plan_execute": [vp, vp, pp, vp, i32],
        "graft_plan_execute_timed": [vp, vp, pp, vp, i32, ctypes.POINTER(ctypes.c_float)],
        "graft_heff_apply": [vp, vp, vp, pp, vp, i32],
        "graft_heff_apply_noalloc": [vp, vp, vp, pp, vp, i32],
        "graft_env_update": [vp, vp, pp, vp],
        "graft_contract_pair": [vp, ctypes.POINTER(PairStepC), ctypes.POINTER(TensorDescC), vp,
                                ctypes.POINTER(TensorDescC), vp, vp],
        "graft_vec_dotc": [vp, i64, vp, vp, dp],
        "graft_vec_nrm2": [vp, i64, vp, dp],
        "graft_vec_axpby": [vp, i64, dp, vp, dp, vp],
        "graft_vec_multi_dotc": [vp, i64, i32, pp, vp, dp],
        "graft_vec_multi_axpy": [vp, i64, i32, pp, dp, vp],
        "graft_vec_multi_dotc_dev": [vp, i64, i32, pp, vp, vp],
        "graft_vec_multi_axpy_dev": [vp, i64, i32, pp, vp, ctypes.c_double, vp],
        "graft_vec_nrm2_dev": [vp, i64, vp, vp],
        "graft_vec_orthonormalize_dev": [vp, i64, i32, pp, vp, ctypes.c_double, vp, vp, vp],
        "graft_vec_scale_inv_dev": [vp, i64, vp, vp],
        "graft_vec_block_combine": [vp, i64, i32, pp, i32, dp, pp],
        "graft_graph_capture_begin": [vp],
        "graft_ctx_reserve_workspace": [vp, ctypes.c_uint64],
        "graft_ctx_set_sector_fusion": [vp, i32],
        "graft_ctx_set_fused_map_loads": [vp, i32],
        "graft_ctx_set_lean_stream": [vp, i32],
        "graft_ctx_set_split_words": [vp, i32],
        "graft_graph_capture_end": [vp, ctypes.POINTER(vp)],
        "graft_graph_launch": [vp, vp],
        "graft_graph_destroy": [vp],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    L.graft_last_error.restype = ctypes.c_char_p
    L.graft_last_error.argtypes = []
    L.graft_version.restype = ctypes.c_char_p
    L.graft_num_tile_configs.restype = ctypes.c_int
    _lib = L
    return L


def check(status):
    if status != 0:
        msg = lib().graft_last_error().decode(errors="replace")
        raise GraftError(status, msg)


def ptr_array(ptrs):
    arr = (ctypes.c_void_p * max(1, len(ptrs)))()
    for i, p in enumerate(ptrs):
        arr[i] = ctypes.c_void_p(int(p))
    return arr


def make_step(step):
    """PairStep (0-based leg tuples) -> PairStepC."""
    s = PairStepC()
    s.a, s.b, s.dst = step.a, step.b, step.dst
    for name in ("oindA", "cindA", "oindB", "cindB"):
        vals = getattr(step, name)
        setattr(s, "n_" + name, len(vals))
        arr = getattr(s, name)
        for i, v in enumerate(vals):
            arr[i] = v
    s.conjA, s.conjB = int(step.conjA), int(step.conjB)
    s.n_out_codomain, s.n_out_domain = len(step.out[0]), len(step.out[1])
    for i, v in enumerate(tuple(step.out[0]) + tuple(step.out[1])):
        s.out[i] = v
    return s


class DescKeepalive:
    """A TensorDescC plus the numpy arrays its pointers reference."""

    def __init__(self, dims, sectors=None):
        self.c = TensorDescC()
        self.c.rank = len(dims)
        for i, d in enumerate(dims):
            self.c.dims[i] = int(d)
            self.c.nsectors[i] = 1
        self.c.nblocks = 1
        self._keep = []
        if sectors is not None:
            secdims, blocks, offsets, strides = sectors
            sd = np.ascontiguousarray(np.concatenate([np.asarray(s, dtype=np.int64) for s in secdims]))
            for i, s in enumerate(secdims):
                self.c.nsectors[i] = len(s)
            bs = np.ascontiguousarray(np.asarray(blocks, dtype=np.int32).reshape(-1))
            bo = np.ascontiguousarray(np.asarray(offsets, dtype=np.int64))
            self._keep += [sd, bs, bo]
            self.c.sector_dims = sd.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
            self.c.nblocks = len(offsets)
            self.c.block_sectors = bs.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
            self.c.block_offsets = bo.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
            if strides is not None:
                st = np.ascontiguousarray(np.asarray(strides, dtype=np.int64).reshape(-1))
                self._keep.append(st)
                self.c.block_strides = st.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


class Context:
    """One graft_ctx: one device, one HIP stream (torch's current stream)."""

    def __init__(self, device=0, stream=None):
        import torch

        L = lib()
        if not torch.cuda.is_available():
            raise RuntimeError("no HIP device visible: the effective-Hamiltonian engine needs a GPU")
        self.device = device
        if stream is None:
            stream = torch.cuda.current_stream(device).cuda_stream
        h = ctypes.c_void_p()
        check(L.graft_ctx_create(device, ctypes.c_void_p(stream), ctypes.byref(h)))
        self.handle = h

    def set_stream(self, stream):
        check(lib().graft_ctx_set_stream(self.handle, ctypes.c_void_p(stream)))

    def set_workspace_cap(self, nbytes):
        check(lib().graft_ctx_set_workspace_cap(self.handle, int(nbytes)))

    def workspace_bytes(self):
        v = ctypes.c_uint64()
        check(lib().graft_ctx_workspace_bytes(self.handle, ctypes.byref(v)))
        return v.value

    def set_sector_fusion(self, mode):
        """0 per-block GEMM items, 1 automatic fused per-charge matrices, 2 always fused."""
        check(lib().graft_ctx_set_sector_fusion(self.handle, int(mode)))

    def set_fused_map_loads(self, on):
        """1 (default): fused FP64 sector steps read and write the blocks through the
        element maps inside the GEMM; 0: gather, GEMM, scatter."""
        check(lib().graft_ctx_set_fused_map_loads(self.handle, int(bool(on))))

    def set_lean_stream(self, on):
        """1 (default): per-block streaming sector steps run the lean kernel where it
        applies; 0: always the general per-block kernel."""
        check(lib().graft_ctx_set_lean_stream(self.handle, int(bool(on))))

    def set_split_words(self, on):
        """1 (default): BF16x3 applies hand split-bf16 words from a streaming step to
        its GEMM; 0: the GEMM converts FP64 operands on every tile load."""
        check(lib().graft_ctx_set_split_words(self.handle, int(bool(on))))

    def reserve_workspace(self, nbytes):
        check(lib().graft_ctx_reserve_workspace(self.handle, int(nbytes)))

    def release_workspace(self):
        check(lib().graft_ctx_release_workspace(self.handle))

    def synchronize(self):
        check(lib().graft_synchronize(self.handle))

    def close(self):
        if getattr(self, "handle", None):
            lib().graft_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def skinny_config():
    """Configuration number of the single-block streaming kernel (after the tiles)."""
    return lib().graft_num_tile_configs() - 2


def block_stream_config():
    """Configuration number of the per-block streaming kernels."""
    return lib().graft_num_tile_configs() - 1


def default_context():
    """Process-wide context on the current device and torch stream."""
    global _default_ctx
    import torch

    if _default_ctx is None:
        _default_ctx = Context(torch.cuda.current_device())
    else:
        cur = torch.cuda.current_stream(_default_ctx.device).cuda_stream
        _default_ctx.set_stream(cur)
    return _default_ctx
