"""ctypes binding of libstile_hip.so (include/stile.h).

Tensors cross the boundary as raw device pointers; the current torch stream is
passed as the HIP stream.  Loading fails loudly when the library is missing:
there is no CPU fallback for any kernel.
"""
import ctypes
import os

import numpy as np

from . import build as _build

STILE_BLK_F32, STILE_BLK_BF16, STILE_BLK_F16 = 0, 1, 2
STILE_RUN_DETERMINISTIC = 1
STILE_NOALLOC_SCRATCH_BYTES = 16 << 20
STILE_OPT_KERNEL, STILE_OPT_LDS_LIST, STILE_OPT_LDS_XCD = 1, 2, 3
_OPTIONS = {"kernel": STILE_OPT_KERNEL, "lds_list": STILE_OPT_LDS_LIST, "lds_xcd": STILE_OPT_LDS_XCD}
_BLK_NAMES = {"float32": STILE_BLK_F32, "fp32": STILE_BLK_F32, "f32": STILE_BLK_F32,
              "bfloat16": STILE_BLK_BF16, "bf16": STILE_BLK_BF16,
              "float16": STILE_BLK_F16, "fp16": STILE_BLK_F16, "f16": STILE_BLK_F16}

_i32p = ctypes.POINTER(ctypes.c_int32)
_f32p = ctypes.POINTER(ctypes.c_float)


class StileEllGroup(ctypes.Structure):
    _fields_ = [("m_blk", ctypes.c_int32), ("nnz_cols", ctypes.c_int32),
                ("num_tiles", ctypes.c_int32), ("real_atomic", ctypes.c_int32),
                ("a", _f32p), ("indices_k", _i32p), ("indices_reordered_i", _i32p), ("a_ids", _i32p)]


class StileTcGroup(ctypes.Structure):
    _fields_ = [("group_size", ctypes.c_int32), ("num_windows", ctypes.c_int32),
                ("nnzb", ctypes.c_int32), ("rows_2d", ctypes.c_int32),
                ("k_notsorted", ctypes.c_int32), ("real_atomic", ctypes.c_int32),
                ("a", _f32p), ("i_indices", _i32p), ("indptr", _i32p), ("indices", _i32p), ("a_ids", _i32p)]


class StileSpmmPack(ctypes.Structure):
    _fields_ = [("M", ctypes.c_int32), ("K", ctypes.c_int32),
                ("num_ell", ctypes.c_int32), ("num_tc", ctypes.c_int32),
                ("ell", ctypes.POINTER(StileEllGroup)), ("tc", ctypes.POINTER(StileTcGroup))]


class StileSddmm1DGroup(ctypes.Structure):
    _fields_ = [("len", ctypes.c_int32), ("mid_idx", _i32p), ("j_idx", _i32p), ("remap", _i32p)]


class StileSddmmTcGroup(ctypes.Structure):
    _fields_ = [("nnzb", ctypes.c_int32), ("blk_k", ctypes.c_int32),
                ("i_indices", _i32p), ("j_indices", _i32p), ("indptr", _i32p),
                ("indices_dense", _i32p), ("indices", _i32p)]


class StileSddmmPack(ctypes.Structure):
    _fields_ = [("M", ctypes.c_int32), ("N", ctypes.c_int32), ("nnz", ctypes.c_int32),
                ("num_1d", ctypes.c_int32), ("num_tc", ctypes.c_int32),
                ("g1d", ctypes.POINTER(StileSddmm1DGroup)),
                ("tc", ctypes.POINTER(StileSddmmTcGroup))]


class StilePlanInfo(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("blk_dtype", ctypes.c_int32),
                ("M", ctypes.c_int32), ("K", ctypes.c_int32),
                ("num_windows", ctypes.c_int32), ("num_chunks", ctypes.c_int32),
                ("block_nnz", ctypes.c_int64), ("sparse_nnz", ctypes.c_int64),
                ("device_bytes", ctypes.c_int64), ("hub_rows", ctypes.c_int32),
                ("num_split", ctypes.c_int32)]


class StileSpmmOptions(ctypes.Structure):
    _fields_ = [("scatter_max_nnz", ctypes.c_int32)]


# Every symbol of include/stile.h with its ctypes signature.
class StileEllBucket(ctypes.Structure):
    _fields_ = [("rows", ctypes.c_int32), ("width", ctypes.c_int32), ("row", ctypes.c_void_p),
                ("cols", ctypes.c_void_p), ("vals", ctypes.c_void_p)]


_SIGNATURES = {
    "stile_csr_spmm_f32": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                          ctypes.c_int64, ctypes.c_void_p]),
    "stile_csr_sddmm_f32": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                           ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                           ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]),
    "stile_csr_row_softmax_f32": (ctypes.c_int, [ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                                                 ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p]),
    "stile_plan_create_spmm": (ctypes.c_int, [ctypes.POINTER(StileSpmmPack), ctypes.c_int32,
                                              ctypes.POINTER(ctypes.c_void_p)]),
    "stile_plan_create_spmm_batched": (ctypes.c_int, [ctypes.POINTER(StileSpmmPack), ctypes.c_int32,
                                                      ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]),
    "stile_plan_create_spmm_ex": (ctypes.c_int, [ctypes.POINTER(StileSpmmPack), ctypes.c_int32, ctypes.c_int32,
                                                 ctypes.POINTER(StileSpmmOptions), ctypes.POINTER(ctypes.c_void_p)]),
    "stile_plan_set_option": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    "stile_plan_create_sddmm": (ctypes.c_int, [ctypes.POINTER(StileSddmmPack), ctypes.c_int32,
                                               ctypes.POINTER(ctypes.c_void_p)]),
    "stile_plan_create_sddmm_batched": (ctypes.c_int, [ctypes.POINTER(StileSddmmPack), ctypes.c_int32,
                                                       ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]),
    "stile_spmm_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32,
                                      ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]),
    "stile_spmm_run_ex": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64,
                                         ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32,
                                         ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]),
    "stile_spmm_workspace_bytes": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                                  ctypes.POINTER(ctypes.c_int64)]),
    "stile_spmm_run_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64,
                                              ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32,
                                              ctypes.c_void_p]),
    "stile_sddmm_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                       ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32,
                                       ctypes.c_void_p, ctypes.c_void_p]),
    "stile_plan_set_values": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]),
    "stile_ell_buckets_spmm_f32": (ctypes.c_int, [ctypes.POINTER(StileEllBucket), ctypes.c_int32, ctypes.c_int32,
                                                  ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                                  ctypes.c_int64, ctypes.c_void_p]),
    "stile_plan_create_attention": (ctypes.c_int, [ctypes.POINTER(StileSpmmPack), ctypes.c_int32,
                                                   ctypes.POINTER(ctypes.c_void_p)]),
    "stile_attention_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_int64, ctypes.c_int32, ctypes.c_float, ctypes.c_void_p,
                                           ctypes.c_int64, ctypes.c_void_p]),
    "stile_plan_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(StilePlanInfo)]),
    "stile_plan_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "stile_last_error": (ctypes.c_char_p, []),
    "stile_abi_version": (ctypes.c_int32, []),
}

_lib = None
# set once this process has made a HIP call through the library (plan
# creation starts the HIP runtime); api.parallel_map then spawns its workers
# instead of forking
_hip_started = False


def hip_started():
    """True once this process has created a device plan (the HIP runtime is
    initialised, so forked children could not use it)."""
    return _hip_started


def _mark_hip():
    global _hip_started
    _hip_started = True


class StileError(RuntimeError):
    pass


def lib_path():
    return _build.LIB_PATH


def load():
    """Load libstile_hip.so (built in-tree by stile_amd.build); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not os.path.exists(path):
        raise StileError(f"{path} is missing: build it with `python -m stile_amd.build` "
                         "(there is no CPU fallback)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc):
    if rc != 0:
        msg = load().stile_last_error().decode(errors="replace")
        raise StileError(f"libstile_hip error {rc}: {msg}")


def _ptr(arr, ctype):
    return arr.ctypes.data_as(ctypes.POINTER(ctype))


def _stream(t=None):
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream(t.device if t is not None else None).cuda_stream)


def blk_dtype_code(dtype):
    if isinstance(dtype, int):
        return dtype
    key = str(dtype).replace("torch.", "")
    if key not in _BLK_NAMES:
        raise ValueError(f"unknown block dtype {dtype!r}")
    return _BLK_NAMES[key]


def _c_i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _c_f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


# ---------------------------------------------------------------------------
# Plain CSR kernels.
# ---------------------------------------------------------------------------

def _dev_check(*ts):
    for t in ts:
        if not t.is_cuda:
            raise StileError("libstile_hip kernels take device tensors")
    if len({t.device for t in ts}) > 1:
        raise StileError("all tensors of a call must be on one device")


def _i32(*ts):
    import torch
    for t in ts:
        if t.dtype != torch.int32 or not t.is_contiguous():
            raise StileError("index arrays must be contiguous int32")


def _out(out, shape, device, what="out"):
    """The output tensor of a call: a new one, or the caller's after checking
    what the kernel will write through its pointer (float32, on `device`,
    exactly `shape`, rows contiguous)."""
    import torch
    if out is None:
        return torch.empty(shape, device=device, dtype=torch.float32)
    if not out.is_cuda or out.device != device:
        raise StileError(f"{what} must be on {device}")
    if out.dtype != torch.float32:
        raise StileError(f"{what} must be float32")
    if tuple(out.shape) != tuple(shape):
        raise StileError(f"{what} has shape {tuple(out.shape)}, expected {tuple(shape)}")
    if out.numel() > 0 and out.stride(-1) != 1:
        raise StileError(f"{what} must have unit stride along its last dimension")
    return out


def csr_spmm_f32(indptr, indices, vals, B, out=None):
    """C = A @ B for CSR A (int32 indptr/indices, float32 vals) on the device."""
    import torch
    _dev_check(indptr, indices, vals, B)
    _i32(indptr, indices)
    M = indptr.numel() - 1
    K, N = B.shape
    if B.dtype != torch.float32 or vals.dtype != torch.float32:
        raise StileError("csr_spmm_f32 takes float32 values and B")
    if vals.numel() != indices.numel() or not vals.is_contiguous():
        raise StileError("vals must be contiguous with one value per index")
    if B.stride(1) != 1:
        B = B.contiguous()
    C = _out(out, (M, N), B.device)
    _check(load().stile_csr_spmm_f32(M, K, N, indptr.data_ptr(), indices.data_ptr(), vals.data_ptr(),
                                     B.data_ptr(), B.stride(0), C.data_ptr(), C.stride(0), _stream(B)))
    return C


def csr_sddmm_f32(indptr, indices, X, Y, out=None):
    """out[p] = <X[row(p)], Y[indices[p]]> over the CSR pattern, on the device."""
    import torch
    _dev_check(indptr, indices, X, Y)
    _i32(indptr, indices)
    M = indptr.numel() - 1
    nnz = indices.numel()
    Kf = X.shape[1]
    if X.dtype != torch.float32 or Y.dtype != torch.float32:
        raise StileError("csr_sddmm_f32 takes float32 X and Y")
    if X.shape[0] < M or Y.shape[1] != Kf:
        raise StileError("X must have a row per CSR row and Y as many features as X")
    if X.stride(1) != 1:
        X = X.contiguous()
    if Y.stride(1) != 1:
        Y = Y.contiguous()
    o = _out(out, (nnz,), X.device)
    _check(load().stile_csr_sddmm_f32(M, Y.shape[0], nnz, Kf, indptr.data_ptr(), indices.data_ptr(),
                                      X.data_ptr(), X.stride(0), Y.data_ptr(), Y.stride(0),
                                      o.data_ptr(), _stream(X)))
    return o


def ell_buckets_spmm_f32(buckets, M, B, C):
    """C = sum of ELL-bucket products (buckets: list of device (row, cols,
    vals, width) tuples as made by stile_amd.hyb.HybSpMM)."""
    import torch
    _dev_check(B, C)
    if B.dtype != torch.float32 or B.stride(1) != 1:
        raise StileError("B must be float32 with contiguous rows")
    _out(C, (M, B.shape[1]), B.device, "C")
    arr = (StileEllBucket * max(len(buckets), 1))(*[
        StileEllBucket(rows=r.shape[0], width=w, row=r.data_ptr(), cols=c.data_ptr(), vals=v.data_ptr())
        for r, c, v, w in buckets])
    _check(load().stile_ell_buckets_spmm_f32(arr, len(buckets), M, B.shape[1], B.data_ptr(), B.stride(0),
                                             C.data_ptr(), C.stride(0), _stream(B)))
    return C


def csr_row_softmax_f32(indptr, vals, scale=1.0, out=None):
    """Softmax of scale*vals over each CSR row (device float32)."""
    import torch
    _dev_check(indptr, vals)
    _i32(indptr)
    if vals.dtype != torch.float32 or not vals.is_contiguous():
        raise StileError("vals must be contiguous float32")
    o = _out(out, tuple(vals.shape), vals.device)
    _check(load().stile_csr_row_softmax_f32(indptr.numel() - 1, indptr.data_ptr(), vals.data_ptr(),
                                            o.data_ptr(), float(scale), _stream(vals)))
    return o


# ---------------------------------------------------------------------------
# Plans.
# ---------------------------------------------------------------------------

class _PackArrays:
    """Keeps the numpy buffers of a ctypes pack alive while the plan is built."""

    def __init__(self):
        self.keep = []

    def i32(self, a):
        a = _c_i32(a)
        self.keep.append(a)
        return _ptr(a, ctypes.c_int32)

    def f32(self, a):
        a = _c_f32(a)
        self.keep.append(a)
        return _ptr(a, ctypes.c_float)

    def i32_opt(self, a):
        return None if a is None else self.i32(a)


def _spmm_pack_struct(pack, arrays):
    """pack: dict with M, K, ell (list of dicts), tc (list of dicts); Appendix-A layouts."""
    ells = [StileEllGroup(m_blk=g["m_blk"], nnz_cols=g["nnz_cols"], num_tiles=g["num_tiles"],
                          real_atomic=int(bool(g.get("real_atomic", False))),
                          a=arrays.f32(g["a"]), indices_k=arrays.i32(g["indices_k"]),
                          indices_reordered_i=arrays.i32(g["indices_reordered_i"]),
                          a_ids=arrays.i32_opt(g.get("a_ids")))
            for g in pack["ell"]]
    tcs = [StileTcGroup(group_size=g["group_size"], num_windows=g["num_windows"], nnzb=g["nnzb"],
                        rows_2d=int(bool(g["rows_2d"])), k_notsorted=int(bool(g.get("k_notsorted", False))),
                        real_atomic=int(bool(g.get("real_atomic", False))),
                        a=arrays.f32(g["a"]), i_indices=arrays.i32(g["i_indices"]),
                        indptr=arrays.i32(g["indptr"]), indices=arrays.i32(g["indices"]),
                        a_ids=arrays.i32_opt(g.get("a_ids")))
           for g in pack["tc"]]
    ell_arr = (StileEllGroup * max(len(ells), 1))(*ells)
    tc_arr = (StileTcGroup * max(len(tcs), 1))(*tcs)
    arrays.keep += [ell_arr, tc_arr]
    return StileSpmmPack(M=pack["M"], K=pack["K"], num_ell=len(ells), num_tc=len(tcs),
                         ell=ctypes.cast(ell_arr, ctypes.POINTER(StileEllGroup)),
                         tc=ctypes.cast(tc_arr, ctypes.POINTER(StileTcGroup)))


def _sddmm_pack_struct(pack, arrays):
    g1 = [StileSddmm1DGroup(len=len(g["mid_idx"]), mid_idx=arrays.i32(g["mid_idx"]),
                            j_idx=arrays.i32(g["j_idx"]), remap=arrays.i32(g["remap"]))
          for g in pack["g1d"]]
    tcs = [StileSddmmTcGroup(nnzb=g["nnzb"], blk_k=g.get("blk_k", 16),
                             i_indices=arrays.i32(g["i_indices"]), j_indices=arrays.i32(g["j_indices"]),
                             indptr=arrays.i32(g["indptr"]), indices_dense=arrays.i32(g["indices_dense"]),
                             indices=arrays.i32(g["indices"]))
           for g in pack["tc"]]
    g1_arr = (StileSddmm1DGroup * max(len(g1), 1))(*g1)
    tc_arr = (StileSddmmTcGroup * max(len(tcs), 1))(*tcs)
    arrays.keep += [g1_arr, tc_arr]
    return StileSddmmPack(M=pack["M"], N=pack["N"], nnz=pack["nnz"], num_1d=len(g1), num_tc=len(tcs),
                          g1d=ctypes.cast(g1_arr, ctypes.POINTER(StileSddmm1DGroup)),
                          tc=ctypes.cast(tc_arr, ctypes.POINTER(StileSddmmTcGroup)))


class Plan:
    """Owner of a StilePlan (device-resident tile buffers of one tile set)."""

    def __init__(self, handle, kind, M, K, nnz=0):
        self._h = handle
        self.kind = kind
        self.M, self.K, self.nnz = M, K, nnz

    @classmethod
    def spmm(cls, packs, blk_dtype="bfloat16", scatter_max_nnz=0):
        """Build from one pack dict or a list of them (block-diagonal batch).
        scatter_max_nnz: B-stationary plans put a window's TC columns with at
        most that many nonzeros into scatter chunks (0 = none)."""
        lib = load()
        packs = packs if isinstance(packs, (list, tuple)) else [packs]
        arrays = _PackArrays()
        structs = (StileSpmmPack * len(packs))(*[_spmm_pack_struct(p, arrays) for p in packs])
        h = ctypes.c_void_p()
        _mark_hip()
        opts = StileSpmmOptions(scatter_max_nnz=int(scatter_max_nnz))
        _check(lib.stile_plan_create_spmm_ex(structs, len(packs), blk_dtype_code(blk_dtype), ctypes.byref(opts),
                                             ctypes.byref(h)))
        plan = cls(h, "spmm", sum(p["M"] for p in packs), sum(p["K"] for p in packs))
        plan.scatter_max_nnz = int(scatter_max_nnz)
        return plan

    def set_option(self, name, value):
        """Launch option of an SpMM plan: 'kernel' (0 B-stationary when the
        plan allows, 1 streaming), 'lds_list' (-1 the plan's pick, 0..11),
        'lds_xcd' (1 per-XCD item ranges, 0 round-robin)."""
        if name not in _OPTIONS:
            raise StileError(f"unknown option {name!r}")
        _check(load().stile_plan_set_option(self._h, _OPTIONS[name], int(value)))

    @classmethod
    def attention(cls, packs):
        """Fused-attention plan over square mask packs (one per instance,
        block-diagonal); also runs as a BF16-block SpMM plan."""
        lib = load()
        packs = packs if isinstance(packs, (list, tuple)) else [packs]
        arrays = _PackArrays()
        structs = (StileSpmmPack * len(packs))(*[_spmm_pack_struct(p, arrays) for p in packs])
        h = ctypes.c_void_p()
        _mark_hip()
        _check(lib.stile_plan_create_attention(structs, len(packs), ctypes.byref(h)))
        plan = cls(h, "spmm", sum(p["M"] for p in packs), sum(p["K"] for p in packs))
        plan.attention_plan = True
        return plan

    def run_attention(self, Q, K, V, scale, out=None):
        """out = softmax(scale * Q K^T on the masks) V; Q, K, V BF16 (rows x 64,
        rows = the stacked instances), out FP32."""
        import torch
        if not getattr(self, "attention_plan", False):
            raise StileError("not an attention plan")
        _dev_check(Q, K, V)
        for t in (Q, K, V):
            if t.dtype != torch.bfloat16 or t.shape != (self.M, 64):
                raise StileError("Q, K, V must be bfloat16 of shape (rows, 64)")
        if not (Q.stride() == K.stride() == V.stride()) or Q.stride(1) != 1:
            Q, K, V = Q.contiguous(), K.contiguous(), V.contiguous()
        o = _out(out, (self.M, 64), Q.device)
        _check(load().stile_attention_run(self._h, Q.data_ptr(), K.data_ptr(), V.data_ptr(), Q.stride(0), 64,
                                          float(scale), o.data_ptr(), o.stride(0), _stream(Q)))
        return o

    @classmethod
    def sddmm(cls, packs, blk_dtype="float32"):
        lib = load()
        packs = packs if isinstance(packs, (list, tuple)) else [packs]
        arrays = _PackArrays()
        structs = (StileSddmmPack * len(packs))(*[_sddmm_pack_struct(p, arrays) for p in packs])
        h = ctypes.c_void_p()
        _mark_hip()
        _check(lib.stile_plan_create_sddmm_batched(structs, len(packs), blk_dtype_code(blk_dtype),
                                                   ctypes.byref(h)))
        return cls(h, "sddmm", sum(p["M"] for p in packs), sum(p["N"] for p in packs),
                   sum(p["nnz"] for p in packs))

    def set_values(self, vals):
        """Refresh A's values (device float32, A's CSR order) without rebuilding."""
        import torch
        if self.kind != "spmm":
            raise StileError("not an SpMM plan")
        _dev_check(vals)
        if vals.dtype != torch.float32:
            raise StileError("values must be float32")
        vals = vals.contiguous()
        _check(load().stile_plan_set_values(self._h, vals.data_ptr(), vals.numel(), _stream(vals)))

    def info(self):
        inf = StilePlanInfo()
        _check(load().stile_plan_info(self._h, ctypes.byref(inf)))
        return {f: getattr(inf, f) for f, _ in StilePlanInfo._fields_}

    def workspace_bytes(self, N, deterministic=False):
        """Device workspace (bytes) run_spmm needs at width N."""
        n = ctypes.c_int64()
        _check(load().stile_spmm_workspace_bytes(self._h, int(N), STILE_RUN_DETERMINISTIC if deterministic else 0,
                                                 ctypes.byref(n)))
        return n.value

    def run_spmm(self, B, out=None, deterministic=False, workspace=None):
        """C = A.B.  B: float32, bfloat16 or float16 (K x N, rows contiguous);
        C float32.  deterministic: split rows and huge windows are summed in a
        fixed order through `workspace` (a device uint8/any tensor of at least
        workspace_bytes(N, True) bytes; allocated here when not given)."""
        import torch
        if self.kind != "spmm":
            raise StileError("not an SpMM plan")
        _dev_check(B)
        codes = {torch.float32: STILE_BLK_F32, torch.bfloat16: STILE_BLK_BF16, torch.float16: STILE_BLK_F16}
        if B.dtype not in codes:
            raise StileError("B must be float32, bfloat16 or float16")
        if B.shape[0] != self.K:
            raise StileError(f"B has {B.shape[0]} rows, the plan's A has {self.K} columns")
        if B.stride(1) != 1:
            B = B.contiguous()
        N = B.shape[1]
        C = _out(out, (self.M, N), B.device)
        flags, ws_ptr, ws_bytes = 0, None, 0
        if deterministic:
            flags = STILE_RUN_DETERMINISTIC
            need = self.workspace_bytes(N, True)
            if need > 0:
                if workspace is None:
                    workspace = torch.empty(need, dtype=torch.uint8, device=B.device)
                _dev_check(workspace)
                if workspace.device != B.device or not workspace.is_contiguous():
                    raise StileError("workspace must be a contiguous tensor on B's device")
                ws_ptr, ws_bytes = workspace.data_ptr(), workspace.numel() * workspace.element_size()
        _check(load().stile_spmm_run_ex(self._h, codes[B.dtype], B.data_ptr(), B.stride(0), N, C.data_ptr(),
                                        C.stride(0), flags, ws_ptr, ws_bytes, _stream(B)))
        return C

    def run_spmm_noalloc(self, B, out, deterministic=False):
        """C = A.B into `out` through stile_spmm_run_noalloc: no allocation,
        the deterministic workspace is the library's static device scratch
        (STILE_NOALLOC_SCRATCH_BYTES per GPU; larger needs are rejected)."""
        import torch
        if self.kind != "spmm":
            raise StileError("not an SpMM plan")
        _dev_check(B, out)
        codes = {torch.float32: STILE_BLK_F32, torch.bfloat16: STILE_BLK_BF16, torch.float16: STILE_BLK_F16}
        if B.dtype not in codes:
            raise StileError("B must be float32, bfloat16 or float16")
        if B.shape[0] != self.K or B.stride(1) != 1:
            raise StileError(f"B must be {self.K} x N with contiguous rows")
        N = B.shape[1]
        if out.dtype != torch.float32 or tuple(out.shape) != (self.M, N) or out.stride(1) != 1:
            raise StileError(f"out must be a float32 {self.M} x {N} tensor with contiguous rows")
        _check(load().stile_spmm_run_noalloc(self._h, codes[B.dtype], B.data_ptr(), B.stride(0), N, out.data_ptr(),
                                             out.stride(0), STILE_RUN_DETERMINISTIC if deterministic else 0,
                                             _stream(B)))
        return out

    def run_sddmm(self, X, Y, out=None):
        import torch
        if self.kind != "sddmm":
            raise StileError("not an SDDMM plan")
        _dev_check(X, Y)
        if X.shape[0] != self.M or Y.shape[0] != self.K or X.shape[1] != Y.shape[1]:
            raise StileError("X/Y shapes do not match the plan")
        if X.dtype != torch.float32 or Y.dtype != torch.float32:
            raise StileError("X and Y must be float32")
        X = X.contiguous() if X.stride(1) != 1 else X
        Y = Y.contiguous() if Y.stride(1) != 1 else Y
        o = _out(out, (self.nnz,), X.device)
        _check(load().stile_sddmm_run(self._h, X.data_ptr(), X.stride(0), Y.data_ptr(), Y.stride(0),
                                      X.shape[1], o.data_ptr(), _stream(X)))
        return o

    def close(self):
        if self._h is not None and self._h.value:
            _check(load().stile_plan_destroy(self._h))
        self._h = None

    def __del__(self):
        try:
            if self._h is not None and self._h.value and _lib is not None:
                _lib.stile_plan_destroy(self._h)
        except Exception:
            pass
        self._h = None
