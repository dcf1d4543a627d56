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
                   precision=_native.PG_FP32, host_params=None, **overrides):
        """Returns (ep, off, ids): the final lists as CSR.  host_params:
        host_stage.make_params overrides (M_nn, rounds, M_final, threads)."""
        n, dim = self.base_.shape
        hp = host_stage.make_params(n, dim, self.gt_ids_.shape[0],
                                    gt_k=self.gt_ids_.shape[1],
                                    **(host_params or {}))
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
