"""Heterogeneous batched forward + adjoint in one persistent, graph-captured launch.

    hb = HeteroBatch([options_0, options_1, ...], device="cuda:0")
    ys, grads = hb.run(xs, ms)      # y_s = A_s x_s,  grad_s = A_s^T (A_s x_s - m_s)

Each slice has its own geometry (image size, views, bins, scan_type).  The
batch owns static device buffers for the inputs and outputs; `run` copies the
inputs in (skip it by writing into `hb.x[s]`, `hb.m[s]` directly and calling
`hb.launch()`), replays a HIP graph holding the counter reset and the one
persistent kernel (ctlib_batch_run), and returns views of the static outputs.
The schedule (resident blocks per CU, item order) comes from the autotuner's
cache when present (ctlib_amd.tune.tune_batch).
"""
import numpy as np
import torch

from . import _C


def _host(o):
    if isinstance(o, torch.Tensor):
        return o.detach().to("cpu", torch.float64)
    return torch.as_tensor(np.asarray(o, dtype=np.float64))


class HeteroBatch:
    def __init__(self, options_list, device="cuda:0", dtype=torch.float32, graph=True, schedule=None):
        self.device = torch.device(device)
        idx = self.device.index if self.device.index is not None else torch.cuda.current_device()
        self.dtype = dtype
        self.options = [_host(o) for o in options_list]
        self.geoms = [_C.parse_options(o) for o in self.options]
        self._b = _C.Batch(self.options, idx, dtype == torch.float64)
        kw = dict(device=self.device, dtype=dtype)
        self.x = [torch.zeros(1, 1, g["height"], g["width"], **kw) for g in self.geoms]
        self.m = [torch.zeros(1, 1, g["views"], g["dets"], **kw) for g in self.geoms]
        self.y = [torch.empty(1, 1, g["views"], g["dets"], **kw) for g in self.geoms]
        self.grad = [torch.empty(1, 1, g["height"], g["width"], **kw) for g in self.geoms]
        self.ws = torch.empty(max(self._b.workspace_bytes(), 1), dtype=torch.uint8, device=self.device)
        self._b.bind(self.x, self.m, self.y, self.grad, self.ws)
        self.use_graph = graph
        self._graph = None
        if schedule is None:
            from . import tune
            schedule = tune.cached_batch_schedule(self.signature())
        if schedule is not None:
            self.set_schedule(*schedule)

    def signature(self):
        """Key of the tuning cache: the multiset of (H, W, V, D, scan_type) plus dtype."""
        keys = sorted((g["height"], g["width"], g["views"], g["dets"], g["scan_type"]) for g in self.geoms)
        return f"{str(self.dtype).split('.')[-1]}|" + ";".join(",".join(map(str, k)) for k in keys)

    def set_schedule(self, blocks_per_cu, order):
        self._b.set_schedule(int(blocks_per_cu), int(order))
        self._graph = None

    def info(self):
        return self._b.info()

    def check(self):
        """Synchronise and raise if the last run flagged a dependency-wait timeout."""
        self._b.check()

    def launch(self):
        """Run the batch on the current stream (graph replay when enabled)."""
        if not self.use_graph:
            self._b.run()
            return
        if self._graph is None:
            self._b.run()  # warm-up outside capture (kernel attributes, code object load)
            torch.cuda.synchronize(self.device)
            s = torch.cuda.Stream(self.device)
            s.wait_stream(torch.cuda.current_stream(self.device))
            g = torch.cuda.CUDAGraph()
            with torch.cuda.stream(s):
                g.capture_begin()
                self._b.run()
                g.capture_end()
            torch.cuda.current_stream(self.device).wait_stream(s)
            self._graph = g
        self._graph.replay()

    def run(self, xs, ms):
        with torch.no_grad():
            for s, (x, m) in enumerate(zip(xs, ms)):
                self.x[s].copy_(x.reshape(self.x[s].shape))
                self.m[s].copy_(m.reshape(self.m[s].shape))
        self.launch()
        return self.y, self.grad
