"""Serving loop: resident point clouds, footprint admission, continuous
batching of render requests over concurrent slots.

    server = PointServer({"sphere": xyz0, "boxes": xyz1}, RenderOptions(...), slots=4,
                         slot_bytes=1 << 30)
    req = server.submit("boxes", cam)          # admission check happens here
    while not req.done:
        server.poll()                          # launches and completes batches
    out = req.result                           # n_samples, sample_t, nb_idx[, nb_t, nb_dist]

The point clouds stay in HBM for the server's lifetime.  Each slot owns a
fixed arena and a stream.  A request's footprint is the arena bytes its view
needs (build workspace, pixel-ordered point list, hash table, per-pixel
bounding spheres and compact outputs); submit() rejects a request that can
never fit a slot.  Whenever a slot is idle, poll() fills it with the oldest
pending requests whose footprints fit its arena (at most max_views), writes
their view descriptors and replays the slot's captured HIP graph.  Batches in
different slots run concurrently on their own streams, so requests that
arrive while others are on the GPU start as soon as a slot frees.  Each
batch's outputs are copied out of the arena on the slot's stream (one copy
per output array) before the slot can be reused; a request's result tensors
are views of that copy, ready when `done` is true, and they equal the
results of the same view sampled alone: every ray is computed independently
of the other views in its batch.  When poll() completes a request it records
the caller's current stream on the result tensors, so their memory is not
reused before work enqueued there has run; a caller that reads them on
another stream records that stream itself (Tensor.record_stream).
"""
import collections
import time

import torch

from .batch import MultiViewSampler


class Request:
    __slots__ = ("id", "cloud", "cam", "arrival", "launched", "finished", "footprint", "rays", "result", "slot")

    def __init__(self, rid, cloud, cam, arrival, footprint, rays):
        self.id = rid
        self.cloud = cloud
        self.cam = cam
        self.arrival = arrival
        self.launched = None
        self.finished = None
        self.footprint = footprint
        self.rays = rays
        self.result = None
        self.slot = None

    @property
    def done(self):
        return self.finished is not None

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival


class _Slot:
    def __init__(self, index, sampler, nbytes):
        self.index = index
        self.sampler = sampler
        self.nbytes = nbytes
        self.requests = []
        self.event = None


class PointServer:
    """Continuous-batching sampler over resident point clouds."""

    def __init__(self, clouds, options, mode="compat", slots=4, slot_bytes=1 << 30, max_views=32, tuning=None,
                 device="cuda:0", clock=time.perf_counter):
        self.device = torch.device(device)
        self.clouds = {}
        for name, xyz in clouds.items():
            if not (xyz.is_cuda and xyz.dtype == torch.float32 and xyz.is_contiguous() and xyz.dim() == 2
                    and xyz.shape[1] == 3):
                raise RuntimeError(f"PointServer: cloud {name!r} must be a contiguous float32 CUDA tensor [n, 3]")
            self.clouds[name] = xyz
        if slots < 1:
            raise RuntimeError("PointServer: at least one slot")
        self.options = options
        self.mode = mode
        self.max_views = int(max_views)
        self.clock = clock
        self.slots = []
        for i in range(int(slots)):
            s = MultiViewSampler(options, mode=mode, device=self.device, max_views=self.max_views, tuning=tuning)
            s.reserve(int(slot_bytes))
            self.slots.append(_Slot(i, s, int(slot_bytes)))
        self.s_max = self.slots[0].sampler.s_max
        self.budget = int(slot_bytes) - self.slots[0].sampler.header_bytes()
        if self.budget <= 0:
            raise RuntimeError("PointServer: slot_bytes too small for the batch header")
        self.pending = collections.deque()
        self.launches = 0
        self.max_inflight = 0  # most slots seen on the GPU at once
        self._next_id = 0
        torch.cuda.synchronize(self.device)  # clouds and arenas ready before any slot stream uses them

    def footprint(self, cloud, cam):
        """Arena bytes a request for `cloud` seen through `cam` occupies."""
        return self.slots[0].sampler.view_bytes(self.clouds[cloud], cam)

    def submit(self, cloud, cam, arrival=None):
        """Queues a request (arrival: its clock time, default now).  Raises
        RuntimeError if it can never fit a slot's arena."""
        if cloud not in self.clouds:
            raise RuntimeError(f"PointServer: unknown cloud {cloud!r}")
        fp = self.footprint(cloud, cam)
        if fp > self.budget:
            raise RuntimeError(f"PointServer: request footprint {fp} B exceeds a slot's arena ({self.budget} B)")
        req = Request(self._next_id, cloud, cam, self.clock() if arrival is None else arrival, fp,
                      int(cam.width) * int(cam.height))
        self._next_id += 1
        self.pending.append(req)
        return req

    def _launch(self, slot):
        take, used = [], 0
        while self.pending and len(take) < self.max_views:
            r = self.pending[0]
            if used + r.footprint > self.budget:
                break
            used += r.footprint
            take.append(self.pending.popleft())
        if not take:
            return False
        s = slot.sampler
        s.set_views([(self.clouds[r.cloud], r.cam) for r in take])
        if not s.captured:
            s.capture()
        now = self.clock()
        with torch.cuda.stream(s.stream):
            s.replay()
            outs, rows = s.batch_outputs()
            outs = {k: v.clone() for k, v in outs.items()}
            ev = torch.cuda.Event()
            ev.record(s.stream)
        for r, (off, R) in zip(take, rows):
            r.launched = now
            r.slot = slot.index
            r.result = {k: v[off:off + R] for k, v in outs.items()}
        slot.requests = take
        slot.event = ev
        self.launches += 1
        self.max_inflight = max(self.max_inflight, sum(sl.event is not None for sl in self.slots))
        return True

    def poll(self):
        """Completes finished batches and starts new ones on idle slots.
        Returns the requests completed by this call."""
        done = []
        for slot in self.slots:
            if slot.event is not None and slot.event.query():
                now = self.clock()
                consumer = torch.cuda.current_stream(self.device)
                for r in slot.requests:
                    r.finished = now
                    # the copies were made on the slot's stream; from here the
                    # caller's stream reads them, so the allocator must wait
                    # for that stream's work before reusing their memory
                    for t in r.result.values():
                        t.record_stream(consumer)
                done.extend(slot.requests)
                slot.requests, slot.event = [], None
        for slot in self.slots:
            if slot.event is None and self.pending:
                self._launch(slot)
        return done

    def busy(self):
        return bool(self.pending) or any(s.event is not None for s in self.slots)

    def drain(self):
        while self.busy():
            if not self.poll():
                time.sleep(1e-5)

    def close(self):
        for s in self.slots:
            s.sampler.close()


def run_trace(server, trace, idle_sleep=2e-5, keep_results=True):
    """Replays a request trace [(t_arrival_s, cloud, cam), ...] (sorted by
    time) on the server's clock and returns the requests.  With
    keep_results=False each result is dropped as soon as its request
    completes (a client that consumed it), so memory stays bounded."""
    t0 = server.clock()
    reqs = []
    i, n = 0, len(trace)
    while i < n or server.busy():
        now = server.clock() - t0
        while i < n and trace[i][0] <= now:
            t, cloud, cam = trace[i]
            reqs.append(server.submit(cloud, cam, arrival=t0 + t))
            i += 1
        done = server.poll()
        if not keep_results:
            for r in done:
                r.result = None
        if not done and not server.pending:
            time.sleep(idle_sleep)
    return reqs


def poisson_trace(rate, count, clouds, seed, resolutions=None, radius=4.0):
    """`count` requests with exponential inter-arrival times of mean 1/rate;
    each picks a cloud uniformly, a resolution uniformly and an eye on the
    sphere of `radius` (deterministic in `seed`)."""
    import numpy as np

    from . import synth
    resolutions = resolutions or synth.RESOLUTIONS
    st = synth.Stream(seed, 7)
    u = st.uniform(count)
    gaps = -np.log(np.maximum(u, 1e-300)) / rate
    times = np.cumsum(gaps) - gaps[0]
    names = sorted(clouds)
    pick = np.minimum((synth.Stream(seed, 8).uniform(count) * len(names)).astype(np.int64), len(names) - 1)
    cams = synth.random_views(count, seed_eye=seed + 1, seed_res=seed + 2, radius=radius, resolutions=resolutions)
    return [(float(times[i]), names[pick[i]], cams[i]) for i in range(count)]
