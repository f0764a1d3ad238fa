"""HIP-graph capture of Monte Carlo Fisher realizations (no counterpart in the reference, whose
compute_fisher runs every realization as a sequence of Python-level calls, PolyBin3D/pspec.py:553-569,
PolyBin3D/bspec.py:678-693).

A pass of a spec's MC Fisher (the realizations of R consecutive seeds; batched for PSpec) is captured
once into a HIP graph (torch.cuda.CUDAGraph on the spec's device) and replayed with new seeds: the
GRFs read their seed on the device (pb3d_grf_device_at), so that a replay differs from the captured
pass only by one device word, which replay() rewrites with a fill kernel before launching the graph.
Every launch of the pass (FFT passes, fused pointing chains, bin reductions, the MFMA contraction, the
FP64 accumulation) is then issued by one graph launch instead of hundreds of host calls, which is what
bounds small grids."""
import gc

import torch


class FisherGraph:
    """replay(s0) adds the Fisher contributions of seeds s0 .. s0 + R - 1 to `out` (an FP64 device tensor
    [N_bins, N_bins]), equal to the eager spec._fisher_contribution_dev of each seed.

    Requires rng='device' and the fused path (trivial S^-1, no covariance).  The constructor runs one
    eager pass (into a scratch matrix) so that every cached work buffer and launch list exists before
    the capture; buffers allocated during the capture belong to the graph's memory pool and are freed
    with it.  Replays on one device are ordered on the current stream; two FisherGraphs of one spec
    share its work buffers, so they must not be replayed on different streams at once."""

    def __init__(self, spec, out, seeds_per_replay=1):
        b = spec.base
        if b.rng != 'device':
            raise ValueError("FisherGraph: the base must draw its GRFs on the device (rng='device')")
        if getattr(spec, 'user_Sinv', False):
            raise ValueError("FisherGraph: a user S^-1 runs host code on every Q map; it cannot be captured")
        if out.dtype != torch.float64 or tuple(out.shape) != (spec.N_bins, spec.N_bins) or out.device != b.device:
            raise ValueError("FisherGraph: out must be an FP64 [N_bins, N_bins] tensor on the spec's device")
        self.spec, self.out, self.R = spec, out, max(1, int(seeds_per_replay))
        self.seed = torch.zeros(1, dtype=torch.int64, device=b.device)
        # the warm-up pass sizes the work buffers that the capture then uses: both run every chain on one
        # stream (a capture stays on the capturing stream)
        self._streams = getattr(spec, 'chain_streams', None)
        if self._streams is not None:
            spec.chain_streams = 1
        try:
            self._capture(spec, b, out)
        finally:
            if self._streams is not None:
                spec.chain_streams = self._streams

    def _capture(self, spec, b, out):
        scratch = torch.zeros_like(out)
        with self._device_seed():
            self._pass(scratch)
        del scratch
        torch.cuda.synchronize(b.device)
        before = {k: v.data_ptr() for k, v in b._bufs.items() if torch.is_tensor(v)}
        self.graph = torch.cuda.CUDAGraph()
        # no garbage collection inside the capture: collecting another graph (or memory it owns) while
        # this stream captures would invalidate the capture
        gc.collect()
        enabled = gc.isenabled()
        gc.disable()
        try:
            with self._device_seed():
                with torch.cuda.graph(self.graph):
                    self._pass(out)
        finally:
            if enabled:
                gc.enable()
        after = {k: v.data_ptr() for k, v in b._bufs.items() if torch.is_tensor(v)}
        if any(after.get(k) != p for k, p in before.items()):
            raise RuntimeError("FisherGraph: a work buffer was reallocated during the capture")

    class _Seed:
        def __init__(self, g):
            self.g = g

        def __enter__(self):
            self.g.spec.base._seed_dev = self.g.seed

        def __exit__(self, *exc):
            self.g.spec.base._seed_dev = None

    def _device_seed(self):
        return FisherGraph._Seed(self)

    def _pass(self, out):
        spec, seeds = self.spec, list(range(self.R))
        if hasattr(spec, 'all_bins'):
            for s in seeds:
                spec._fisher_contribution_dev(s, out=out)
        elif len(seeds) > 1:
            out.add_(spec._fisher_contributions_dev(seeds))
        else:
            out.add_(spec._fisher_contribution_dev(seeds[0]))

    def replay(self, seed0):
        """Add the contributions of seeds seed0 .. seed0 + R - 1 to out (asynchronous, current stream)."""
        self.seed.fill_(int(seed0))
        self.graph.replay()
