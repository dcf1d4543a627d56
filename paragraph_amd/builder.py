"""Builder: the reference's (commented-out) Gbuilder::Builder<T> interface.

Reference stub, include/Builder.hpp:6-41:
    explicit Builder(std::vector<T> const& base, std::vector<id_type> const& gt_ids)
    template <grid_size, block_size, gt_num, dim> auto LaunchBuild() -> void
    auto SetGt(std::vector<id_type>&& gt_ids) -> void
Same constructor arguments and SetGt; LaunchBuild takes the build knobs at
run time (the reference bakes them into template arguments,
tests/test_t2iend.cu:670-727) and returns the two arrays the reference copies
back to the host: the projection rows (N x 16) and the link rows (N x M).

BuildIndex() is the reference's whole build (tests/test_t2iend.cu:850-1258):
the GPU stage, with the host bipartite stage (MatchSup rounds + FusionNN)
running on a CPU thread beside it, then FusionFinal, the entry point and,
given a path, the index file in the SaveGraph layout.
"""
import threading

import numpy as np

from . import _native, host_stage


def resolve_host_stages(host_stages, n_base, n_query):
    """Where BuildIndex runs the host stages: "gpu" needs n_query ==
    n_base (pg_run_fusion indexes its supply lists by base row)."""
    if host_stages == "auto":
        return "gpu" if n_query == n_base else "host"
    if host_stages not in ("gpu", "host"):
        raise ValueError('host_stages must be "auto", "gpu" or "host"')
    if host_stages == "gpu" and n_query != n_base:
        raise ValueError("host_stages='gpu' needs one gt row per base row "
                         f"({n_query} gt rows, {n_base} base rows); use "
                         "'host' or 'auto'")
    return host_stages


class Builder:
    def __init__(self, base, gt_ids, dim=None):
        base = np.ascontiguousarray(base, dtype=np.float32)
        if base.ndim == 1:
            if dim is None:
                raise ValueError("flat base needs dim")
            base = base.reshape(-1, dim)
        self.base_ = base
        self.gt_ids_ = None
        self.SetGt(gt_ids)
        self.stats_ = None

    def SetGt(self, gt_ids, gt_k=128):
        g = np.asarray(gt_ids)
        if g.ndim == 1:
            g = g.reshape(-1, gt_k)
        self.gt_ids_ = np.ascontiguousarray(g.astype(np.int64) & 0xFFFFFFFF,
                                            dtype=np.uint32)

    def LaunchBuild(self, M=32, L=128, device=0, precision=_native.PG_FP32,
                    **overrides):
        if isinstance(precision, str):
            names = {"fp32": _native.PG_FP32, "bf16": _native.PG_BF16,
                     "fp8": _native.PG_FP8E4M3}
            if precision not in names:
                raise ValueError(f"precision must be one of {sorted(names)}")
            precision = names[precision]
        n, dim = self.base_.shape
        q, gt_k = self.gt_ids_.shape
        p = _native.make_params(n, dim, q, M, L, gt_k=gt_k,
                                precision=precision, **overrides)
        proj = np.empty((n, p.proj_deg), np.uint32)
        link = np.empty((n, p.M), np.uint32)
        with _native.Context(p, device) as ctx:
            ctx.stage_host(self.base_, self.gt_ids_, proj, link)
            ctx.wait()
            self.stats_ = ctx.stats().as_dict()
        return proj, link

    def BuildIndex(self, M=32, L=128, path=None, device=0,
                   precision=_native.PG_FP32, host_params=None,
                   host_stages="auto", **overrides):
        """Returns (ep, off, ids): the final lists as CSR.  host_params:
        host_stage.make_params overrides (M_nn, rounds, M_final, threads).
        host_stages="gpu" runs MatchSup / FusionNN / FusionFinal on the GPU
        after the GPU stage (pg_run_fusion), which needs one gt row per
        base row; "host" runs them on the host cores, the bipartite part
        beside the GPU stage as in the reference, for any gt row count;
        "auto" (the default) takes "gpu" when the gt has one row per base
        row and "host" otherwise.  Both give the same lists."""
        n, dim = self.base_.shape
        hp = host_stage.make_params(n, dim, self.gt_ids_.shape[0],
                                    gt_k=self.gt_ids_.shape[1],
                                    **(host_params or {}))
        host_stages = resolve_host_stages(host_stages, n,
                                          self.gt_ids_.shape[0])
        if host_stages == "gpu":
            return self._build_index_gpu(hp, M, L, path, device, precision,
                                         overrides)
        bip = {}

        def host_work():
            try:
                bip["csr"] = host_stage.bipartite(hp, self.base_, self.gt_ids_)
            except Exception as e:  # re-raised on the caller's thread
                bip["err"] = e

        # the bipartite stage overlaps the GPU stage, as in the reference
        # (a CPU thread beside compute_stream, tests/test_t2iend.cu:888-962)
        t = threading.Thread(target=host_work)
        t.start()
        try:
            proj, link = self.LaunchBuild(M=M, L=L, device=device,
                                          precision=precision, **overrides)
        finally:
            t.join()
        if "err" in bip:
            raise bip["err"]
        ep = host_stage.entry_point(hp, self.gt_ids_)
        off, ids = host_stage.fusion_final(hp, self.base_, *bip["csr"], link,
                                           proj)
        if path is not None:
            host_stage.save_index(path, ep, off, ids)
        return ep, off, ids

    def _build_index_gpu(self, hp, M, L, path, device, precision, overrides):
        import torch
        if isinstance(precision, str):
            precision = {"fp32": _native.PG_FP32, "bf16": _native.PG_BF16,
                         "fp8": _native.PG_FP8E4M3}[precision]
        n, dim = self.base_.shape
        q, gt_k = self.gt_ids_.shape
        p = _native.make_params(n, dim, q, M, L, gt_k=gt_k,
                                precision=precision, **overrides)
        dev = torch.device(f"cuda:{device}")
        base_d = torch.from_numpy(self.base_).to(dev)
        gt_d = torch.from_numpy(self.gt_ids_.view(np.int32)).to(dev)
        i32 = dict(dtype=torch.int32, device=dev)
        proj_d = torch.empty((n, p.proj_deg), **i32)
        link_d = torch.empty((n, p.M), **i32)
        width = int(host_stage.lib().pg_host_final_width(
            host_stage.ctypes.byref(hp)))
        off_d = torch.empty(n + 1, **i32)
        ids_d = torch.empty(max(n * width, 1), **i32)
        eps = {}
        # the entry-point scan (sequential) runs on the host meanwhile
        t = threading.Thread(target=lambda: eps.setdefault(
            "ep", host_stage.entry_point(hp, self.gt_ids_)))
        t.start()
        try:
            with _native.Context(p, device) as ctx:
                ctx.stage_device(base_d, gt_d, proj_d, link_d)
                ctx.run_fusion(hp, base_d, gt_d, proj_d, link_d, off_d,
                               ids_d)
                ctx.wait()
                self.stats_ = ctx.stats().as_dict()
        finally:
            t.join()
        off = off_d.cpu().numpy().view(np.uint32)
        ids = ids_d[:int(off[-1])].cpu().numpy().view(np.uint32).copy()
        ep = eps["ep"]
        if path is not None:
            host_stage.save_index(path, ep, off, ids)
        return ep, off, ids
