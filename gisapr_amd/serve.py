"""Candidate-evaluation serving with continuous batching over resident datasets.

Optimizer runs (PSO swarms, simplex vertices, grid points) send requests of
1-64 rigid-body candidates against one of several datasets of different box
sizes.  The server keeps each dataset's particle psi spectra (or their
in-disk coefficients), fixed-body density and direction tables resident in
HBM, admitted against an HBM budget, and runs a continuous-batching loop:

* a step of a dataset takes up to ``slots`` pending candidates from its
  oldest requests (a request larger than the step is spread over several
  steps; requests that arrive while steps are running join the next one);
* every dataset has two executors (``CandidateBatch`` buffers) on two HIP
  streams, so one step's host staging and the next step's launches overlap
  the running step, and steps of different datasets overlap on the device;
* a request completes when its last candidate's score has been copied back;
  latency = completion - arrival.

The worker protocol this extends is gisapr_worker.py:39-173 /
gisapr_runtime.py:153-278 (one evaluation per request); per-candidate
results are the conformation score sum_b exp(rawCC_b / 1024)
(new_method_to_fit_the_2nd_Gaussian_PEAK_v4.py:133-134).

``Scheduler`` is the host policy (no device calls; unit-tested on the CPU);
``SearchServer`` binds it to the device executors.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field
from typing import Callable, Deque, Dict, List, Optional, Sequence, Tuple

import numpy as np


@dataclass
class Request:
    rid: int
    dataset: str
    poses: np.ndarray                 # [k, 6] rotvec (deg) + translation (A)
    arrival: float                    # seconds on the server clock
    scores: Optional[np.ndarray] = None
    scheduled: int = 0                # poses handed to steps so far
    finished: int = 0                 # poses whose scores are back
    t_done: Optional[float] = None

    def __post_init__(self):
        self.poses = np.asarray(self.poses, np.float64).reshape(-1, 6)
        if len(self.poses) == 0:
            raise ValueError("a request needs at least one candidate")
        self.scores = np.full(len(self.poses), np.nan)

    @property
    def size(self) -> int:
        return len(self.poses)

    @property
    def latency(self) -> Optional[float]:
        return None if self.t_done is None else self.t_done - self.arrival


WorkItem = Tuple[Request, int]        # (request, candidate index)


class Scheduler:
    """FIFO continuous batching per dataset.

    ``arrive(req)`` queues a request; ``take(dataset, slots)`` returns up to
    ``slots`` (request, candidate) items from the oldest requests of that
    dataset; ``finish(items, scores, now)`` records scores and completes
    requests.  Pure host logic."""

    def __init__(self):
        self.queues: Dict[str, Deque[Request]] = {}
        self.completed: List[Request] = []

    def arrive(self, req: Request):
        self.queues.setdefault(req.dataset, deque()).append(req)

    def pending(self, dataset: str) -> int:
        return sum(r.size - r.scheduled for r in self.queues.get(dataset, ()))

    def take(self, dataset: str, slots: int) -> List[WorkItem]:
        items: List[WorkItem] = []
        for req in self.queues.get(dataset, ()):
            while req.scheduled < req.size and len(items) < slots:
                items.append((req, req.scheduled))
                req.scheduled += 1
            if len(items) == slots:
                break
        return items

    def finish(self, items: Sequence[WorkItem], scores: Sequence[float], now: float):
        for (req, j), s in zip(items, scores):
            req.scores[j] = s
            req.finished += 1
            if req.finished == req.size:
                req.t_done = now
                self.completed.append(req)
        for q in self.queues.values():
            while q and q[0].finished == q[0].size:
                q.popleft()

    def idle(self) -> bool:
        return all(not q for q in self.queues.values())


def latency_stats(reqs: Sequence[Request]) -> dict:
    lat = np.asarray([r.latency for r in reqs if r.latency is not None], np.float64) * 1000.0
    if len(lat) == 0:
        return dict(n=0, p50_ms=None, p99_ms=None, mean_ms=None, max_ms=None)
    return dict(n=int(len(lat)), p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                mean_ms=float(lat.mean()), max_ms=float(lat.max()))


def queue_growth(reqs: Sequence[Request], wall: float) -> dict:
    """Whether the server kept up with the stream: the requests (and their
    candidates) still open when the last one arrived, the time to drain
    them, and the slope of latency against arrival time (about 0 when the
    server keeps up; 1 - capacity / offered load when the queue grows)."""
    if not reqs:
        return {}
    last = max(r.arrival for r in reqs)
    open_reqs = [r for r in reqs if r.t_done is not None and r.t_done > last]
    arr = np.asarray([r.arrival for r in reqs], np.float64)
    lat = np.asarray([r.latency for r in reqs], np.float64)
    slope = float(np.polyfit(arr, lat, 1)[0]) if len(reqs) > 2 and np.ptp(arr) > 0 else 0.0
    return dict(last_arrival_s=float(last), open_requests_at_last_arrival=len(open_reqs),
                open_candidates_at_last_arrival=int(sum(r.size for r in open_reqs)),
                drain_s=float(wall - last), latency_growth_s_per_s=slope)


def poisson_requests(n: int, rate: float, datasets: Sequence[str], seed: int = 0, mean_candidates: float = 8.0,
                     max_candidates: int = 64, pose_sigma=(2.0, 0.7), weights=None) -> List[Request]:
    """Seeded stream: Poisson arrivals at ``rate``/s, 1..max_candidates candidates
    per request (geometric with the given mean, truncated), random dataset."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    p = 1.0 / float(mean_candidates)
    k = np.minimum(rng.geometric(p, n), max_candidates)
    ds = rng.choice(len(datasets), n, p=weights)
    out = []
    for i in range(n):
        rot = rng.normal(0.0, pose_sigma[0], (k[i], 3))
        tr = rng.normal(0.0, pose_sigma[1], (k[i], 3))
        out.append(Request(i, datasets[ds[i]], np.concatenate([rot, tr], 1), float(t[i])))
    return out


# ---------------------------------------------------------------------------
# Device side
# ---------------------------------------------------------------------------

@dataclass
class Dataset:
    name: str
    engine: object                    # SearchEngine
    particles: object                 # ParticleSet (prepared)
    body: object                      # batch.RigidBody
    res: float
    executors: list = field(default_factory=list)
    streams: list = field(default_factory=list)
    resident_bytes: int = 0


class SearchServer:
    """Continuous-batching evaluation of candidate requests on one GPU.

    ``add_dataset`` makes a dataset resident if its particle data, fixed
    density and two executors fit the remaining HBM budget (else raises
    MemoryError).  ``serve(requests)`` replays a request stream against the
    wall clock (each request becomes visible at its arrival time) and
    returns the completed requests and statistics."""

    def __init__(self, device=None, hbm_budget_bytes: Optional[int] = None, path: str = "mixed", slots: int = 16,
                 autotuner=None):
        import torch
        self.torch = torch
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        if hbm_budget_bytes is None:
            free, total = torch.cuda.mem_get_info(self.device)
            hbm_budget_bytes = int(free * 0.9)
        self.budget = int(hbm_budget_bytes)
        self.used = 0
        self.path = path
        self.slots = int(slots)
        self.autotuner = autotuner
        self.datasets: Dict[str, Dataset] = {}
        self.scheduler = Scheduler()
        self._inflight: Dict[Tuple[str, int], tuple] = {}  # (dataset, executor) -> (items, host scores, event)
        self._reported = 0
        self.pairs = 0                    # (direction, psi) pairs searched so far
        self.steps = 0

    def add_dataset(self, name: str, engine, particles, body, res: float) -> Dataset:
        from .batch import CandidateBatch
        torch = self.torch
        if name in self.datasets:
            raise ValueError(f"dataset {name!r} already resident")
        before = torch.cuda.memory_allocated(self.device)
        ds = Dataset(name, engine, particles, body, float(res))
        if self.path == "mixed":
            if getattr(engine, "small", None) is None:
                engine.enable_small_shift()
            if particles.coeffs is None:
                engine.prepare_small_shift(particles, drop_spectra=True)
        elif particles.spectra_frag is None:
            engine.prepare(particles)
        for _ in range(2):
            ds.executors.append(CandidateBatch(engine, particles, body, res, self.slots, path=self.path,
                                               autotuner=self.autotuner))
            ds.streams.append(torch.cuda.Stream(device=self.device))
        torch.cuda.synchronize(self.device)
        ds.resident_bytes = max(0, torch.cuda.memory_allocated(self.device) - before) + particles.nbytes_resident()
        if self.used + ds.resident_bytes > self.budget:
            ds.executors.clear()
            raise MemoryError(f"dataset {name!r} needs {ds.resident_bytes / 2**30:.2f} GiB, "
                              f"{(self.budget - self.used) / 2**30:.2f} GiB of the HBM budget left")
        self.used += ds.resident_bytes
        self.datasets[name] = ds
        return ds

    # -- incremental loop (the worker drives it between messages) --------------
    def submit(self, req: Request) -> None:
        """Queue a request (its ``arrival`` on the caller's clock)."""
        if req.dataset not in self.datasets:
            raise KeyError(f"request {req.rid}: dataset {req.dataset!r} is not resident")
        self.scheduler.arrive(req)

    def busy(self) -> bool:
        return not self.scheduler.idle() or bool(self._inflight)

    def pump(self, now: Callable[[], float]) -> bool:
        """Retire finished steps, then launch a step on every idle executor
        whose dataset has pending candidates.  ``now()`` stamps completions.
        Returns whether a step was launched."""
        torch = self.torch
        sched = self.scheduler
        for key in list(self._inflight):
            items, host, ev = self._inflight[key]
            if ev.query():
                sched.finish(items, host[:len(items)].tolist(), now())
                del self._inflight[key]
        launched = False
        for name, ds in self.datasets.items():
            for e, ex in enumerate(ds.executors):
                if (name, e) in self._inflight or sched.pending(name) == 0:
                    continue
                items = sched.take(name, self.slots)
                poses = np.stack([req.poses[j] for req, j in items])
                with torch.cuda.stream(ds.streams[e]):
                    scores, corr = ex.evaluate(poses, sync=False)
                    host = torch.empty(scores.shape, dtype=scores.dtype, pin_memory=True)
                    host.copy_(scores, non_blocking=True)
                    ev = torch.cuda.Event()
                    ev.record()
                self._inflight[(name, e)] = (items, host, ev)
                self.pairs += ex.plan.n_pairs * len(items) // ex.slots
                self.steps += 1
                launched = True
        return launched

    def take_completed(self) -> List[Request]:
        """Requests completed since the last call, in completion order."""
        done = self.scheduler.completed[self._reported:]
        self._reported = len(self.scheduler.completed)
        return done

    def serve(self, requests: Sequence[Request], clock: Callable[[], float] = time.perf_counter,
              poll_s: float = 0.0002) -> dict:
        """Replay ``requests`` against the wall clock (each visible from its
        arrival time) until all complete; returns latency and throughput."""
        pending = sorted(requests, key=lambda r: r.arrival)
        for r in pending:
            if r.dataset not in self.datasets:
                raise KeyError(f"request {r.rid}: dataset {r.dataset!r} is not resident")
        t0 = clock()
        rel = lambda: clock() - t0
        first_done = len(self.scheduler.completed)
        pairs0, steps0 = self.pairs, self.steps
        nxt = 0
        while nxt < len(pending) or self.busy():
            now = rel()
            while nxt < len(pending) and pending[nxt].arrival <= now:
                self.submit(pending[nxt])
                nxt += 1
            launched = self.pump(rel)
            if not launched and self._inflight:
                time.sleep(poll_s)
            elif not launched and nxt < len(pending):
                time.sleep(max(0.0, min(pending[nxt].arrival - rel(), 0.002)))
        wall = rel()
        done = self.scheduler.completed[first_done:]
        self._reported = len(self.scheduler.completed)
        stats = latency_stats(done)
        cands = int(sum(r.size for r in done))
        pairs, steps = self.pairs - pairs0, self.steps - steps0
        stats.update(wall_s=wall, requests=len(done), candidates=cands, steps=steps,
                     candidates_per_s=cands / wall if wall > 0 else None,
                     scores_per_s=pairs / wall if wall > 0 else None)
        stats.update(queue_growth(done, wall))
        return stats
