"""Batched sparse attention over many uneven masks in one graph launch.

For instance b with mask S_b (L_b x L_b), Q_b, K_b, V_b (L_b x d):
    scores = SDDMM(Q_b, K_b) on S_b           (hybrid SDDMM plan)
    P      = row_softmax(scores / sqrt(d))    (stile_csr_row_softmax_f32)
    out_b  = P . V_b                          (hybrid SpMM plan, values = P)
All instances are one block-diagonal problem: one SDDMM plan and one SpMM
plan over the stacked rows (stile_plan_create_*_batched), so a forward is
four kernel launches (SDDMM, softmax, value refresh, SpMM) regardless of the
number of instances, and the whole forward can be captured into a HIP graph
(torch.cuda.CUDAGraph) and replayed.

The reference has the masks only as SpMM test patterns (test_LogSparse /
test_Strided, gen_formats_v2.py:2926-3003) and batches only one pattern over
many B (my_fuse_formats.py:1296-1428); this pipeline is the batched-attention
workload built from its SDDMM and SpMM.
"""
import math
import os
import numpy as np
import scipy.sparse
import torch

from . import _native, api


def _pack_one(args):
    kind, A, opts = args
    ones = scipy.sparse.csr_matrix((np.ones(A.nnz, np.float32), A.indices, A.indptr), shape=A.shape)
    if kind == "spmm":
        return api.spmm_pack(ones, opts)[0]
    return api.sddmm_pack(ones, opts)[0]


def build_packs(kind, mats, opts=None, workers=None):
    """Search + pack every matrix, in parallel worker processes."""
    return api.parallel_map(_pack_one, [(kind, A, opts) for A in mats], workers)


def _tune_attention(ones, opts, workers, log, cache_dir, device):
    """Fused-attention plan: the search setting (see api.tuning_candidates)
    whose plan runs the fused kernel fastest on random Q, K, V."""
    import time as _time
    cache = None if cache_dir is None else os.path.join(cache_dir, "attention_fused.npz")
    fp = api.matrix_fingerprint("attention", ones)
    if cache is not None and os.path.exists(cache):
        return _native.Plan.attention(api.load_pack(cache, expect=fp)), {"choice": "cached", "cache": cache}
    cands = api.tuning_candidates("bfloat16")
    packs, search_s = {}, {}
    for name, o in cands.items():
        t0 = _time.perf_counter()
        packs[name] = build_packs("spmm", ones, o, workers)
        search_s[name] = round(_time.perf_counter() - t0, 3)
        if log:
            log(f"tune attention: searched {name} in {search_s[name]} s")
    n = sum(m.shape[0] for m in ones)
    g = torch.Generator(device=device)
    g.manual_seed(0)
    Q, K, V = (torch.randn(n, 64, device=device, generator=g).to(torch.bfloat16) for _ in range(3))
    best, best_ms, times, choice = None, float("inf"), {}, None
    for name, pk in packs.items():
        plan = _native.Plan.attention(pk)
        for _ in range(3):
            plan.run_attention(Q, K, V, 0.125)
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(10):
            plan.run_attention(Q, K, V, 0.125)
        e.record()
        torch.cuda.synchronize()
        ms = s.elapsed_time(e) / 10
        times[name] = round(ms, 5)
        if log:
            log(f"tune attention {name}: {ms * 1e3:.1f} us")
        if ms < best_ms:
            if best is not None:
                best.close()
            best, best_ms, choice = plan, ms, name
        else:
            plan.close()
    if cache is not None:
        os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
        api.save_pack(cache, packs[choice], fingerprint=fp)
    return best, {"choice": choice, "ms": times, "search_s": search_s}


def _cached_packs(kind, mats, opts, workers, cache_dir, name):
    """build_packs, read from / written to cache_dir/name when cache_dir is
    given (with the masks' fingerprint), so a cached run starts no search."""
    path = None if cache_dir is None else os.path.join(cache_dir, name)
    fp = api.matrix_fingerprint(kind, mats)
    if path is not None and os.path.exists(path):
        return api.load_pack(path, expect=fp)
    packs = build_packs(kind, mats, opts, workers)
    if path is not None:
        os.makedirs(cache_dir, exist_ok=True)
        api.save_pack(path, packs, fingerprint=fp)
    return packs


class BatchedSparseAttention(object):
    """Sparse attention over a list of CSR masks (one per instance)."""

    def __init__(self, masks, head_dim, blk_dtype="bfloat16", opts=None, device="cuda", workers=None,
                 tune=False, log=None, cache_dir=None, fused=None):
        self.masks = [scipy.sparse.csr_matrix(m) for m in masks]
        for m in self.masks:
            m.sort_indices()
        self.lens = [m.shape[0] for m in self.masks]
        self.d = head_dim
        self.device = torch.device(device)
        self.row_off = np.concatenate([[0], np.cumsum(self.lens)])
        self.tuning = None
        # fused: one kernel (stile_attention_run) for head dimension 64;
        # otherwise SDDMM -> row softmax -> value refresh -> SpMM
        self.fused = (head_dim == 64) if fused is None else fused
        if self.fused:
            if head_dim != 64:
                raise ValueError("the fused kernel supports head dimension 64")
            ones = [scipy.sparse.csr_matrix((np.ones(m.nnz, np.float32), m.indices, m.indptr), shape=m.shape)
                    for m in self.masks]
            if tune:
                self.att_plan, self.tuning = _tune_attention(ones, opts, workers, log, cache_dir, device)
            else:
                packs = _cached_packs("spmm", self.masks, api.options_for(blk_dtype, opts), workers, cache_dir,
                                      "attention_fused_default.npz")
                self.att_plan = _native.Plan.attention(packs)
        elif tune:
            # one search setting for all instances, chosen by measured latency
            ones = [scipy.sparse.csr_matrix((np.ones(m.nnz, np.float32), m.indices, m.indptr), shape=m.shape)
                    for m in self.masks]
            cp = (lambda name: None if cache_dir is None else os.path.join(cache_dir, name))
            self.spmm_plan, _, rep = api.tune_plan("spmm", ones, blk_dtype, head_dim, workers=workers, log=log,
                                                   cache=cp("spmm.npz"))
            self.sddmm_plan, _, srep = api.tune_plan("sddmm", ones, "float32", head_dim, workers=workers, log=log,
                                                     cache=cp("sddmm.npz"))
            self.tuning = {"spmm": rep, "sddmm": srep}
        else:
            spmm_packs = _cached_packs("spmm", self.masks, api.options_for(blk_dtype, opts), workers, cache_dir,
                                       "unfused_spmm.npz")
            sddmm_packs = _cached_packs("sddmm", self.masks, api.options_for("float32", opts), workers, cache_dir,
                                        "unfused_sddmm.npz")
            self.spmm_plan = _native.Plan.spmm(spmm_packs, blk_dtype)
            self.sddmm_plan = _native.Plan.sddmm(sddmm_packs, "float32")
        big = scipy.sparse.block_diag(self.masks, format="csr")
        big.sort_indices()
        self.nnz = big.nnz
        self.indptr = torch.from_numpy(big.indptr.astype(np.int32)).to(self.device)
        self.indices = big.indices.astype(np.int32)
        self.scale = 1.0 / math.sqrt(head_dim)
        if not self.fused:
            self.scores = torch.empty(self.nnz, device=self.device, dtype=torch.float32)
            self.probs = torch.empty(self.nnz, device=self.device, dtype=torch.float32)
        self.graph = None

    def forward(self, Q, K, V, out=None):
        """Q, K, V: (sum L_b) x d stacked.  Fused: BF16 operands (others are
        converted), one kernel; unfused: computed in FP32."""
        if self.fused:
            q, k, v = (t if t.dtype == torch.bfloat16 else t.to(torch.bfloat16) for t in (Q, K, V))
            return self.att_plan.run_attention(q, k, v, self.scale, out=out)
        q = Q.float().contiguous()
        k = K.float().contiguous()
        v = V.float().contiguous()
        self.sddmm_plan.run_sddmm(q, k, out=self.scores)
        _native.csr_row_softmax_f32(self.indptr, self.scores, self.scale, out=self.probs)
        self.spmm_plan.set_values(self.probs)
        return self.spmm_plan.run_spmm(v, out=out)

    def capture(self, Q, K, V):
        """Capture forward(Q, K, V) into a HIP graph; replay() reruns it on the
        same input/output buffers (update them in place between replays)."""
        self.gQ, self.gK, self.gV = Q, K, V
        self.gout = torch.empty((Q.shape[0], self.d), device=self.device, dtype=torch.float32)
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            for _ in range(2):
                self.forward(Q, K, V, out=self.gout)
        torch.cuda.current_stream().wait_stream(s)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.forward(self.gQ, self.gK, self.gV, out=self.gout)
        return self.gout

    def replay(self):
        self.graph.replay()
        return self.gout
