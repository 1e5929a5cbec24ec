"""Persistent GPU worker: one process per GPU that imports PyTorch and the HIP
library once and then executes requests from a controller over an
authenticated loopback connection.

The request protocol is the reference worker's (gisapr_worker.py:39-173; the
zip release's ZL gisapr_worker_v2.py:120-139 adds ``search_batch_from_pdb``):
each request is a dict with a ``command`` and an optional ``cwd`` (the job's
working directory; the worker's starting directory otherwise); each reply is
``{"ok": True, "result": ...}`` or ``{"ok": False, "error": "<Type>: <msg>",
"traceback": ...}``, and a failed request (argparse's SystemExit included)
leaves the worker running.

  geometry               geometry_kwargs, pose, optional pivot -> (overlap, distance)
  generate_mrc           mrc_kwargs -> generate_mrc's return value
  search                 argv (the search program's arguments) -> run_search metrics
  search_from_pdb        argv, model_kwargs: generate_volume on the worker's
                         GPU, then run_search on that volume (density_s added)
  search_batch_from_pdb  argv, options, models, optional orientation_rows /
                         calibrate -> BatchSearchEngine.run results
  ping                   -> pid, device and resident-cache sizes
  shutdown               -> None, then the worker exits

Serving (continuous batching over resident datasets; no counterpart in the
reference, which evaluates one request at a time):

  add_dataset            name, argv (particles, angles and search settings),
                         model_kwargs (source_pdb, chain_id, res, optional
                         include_hetatm), options (path, slots, hbm_budget_mb)
                         -> resident bytes; refused (MemoryError) when the
                         dataset does not fit the HBM budget
  evaluate               rid, dataset, poses [k][6] -> answered when complete
                         as {"ok": True, "rid": rid, "result": {"scores",
                         "latency_s"}}; requests that arrive while steps run
                         join the next steps, so replies can come out of order

The worker's GPU overrides any per-request device.  ``--worker-gpuid cpu``
starts a worker without a GPU: the transport answers (ping, shutdown), and
every compute command fails with the library's "HIP devices only" error.
"""
from __future__ import annotations

import argparse
import os
import sys
import time
import traceback
from collections import OrderedDict
from pathlib import Path

LOOPBACK = ("127.0.0.1", "localhost")


def create_worker_parser() -> argparse.ArgumentParser:
    p = argparse.ArgumentParser(description="GisAPR HIP worker")
    p.add_argument("--worker-connect", nargs=2, metavar=("HOST", "PORT"), required=True)
    p.add_argument("--worker-auth", required=True, help="connection key, hexadecimal")
    p.add_argument("--worker-gpuid", default="0", help="GPU index, or cpu")
    return p


def geometry_key(kwargs: dict) -> tuple:
    """Identity of a prepared GeometryRestraint: its settings plus the PDB
    file's resolved path, size and modification / change times, so an edited
    file is prepared again."""
    path = Path(kwargs["pdb_path"]).expanduser().resolve()
    info = path.stat()
    settings = dict(kwargs, pdb_path=str(path))
    return tuple(sorted(settings.items())), info.st_size, info.st_mtime_ns, info.st_ctime_ns


def _parse_gpu(text) -> object:
    gpu = str(text).strip().lower()
    if gpu == "cpu":
        return "cpu"
    value = int(gpu)
    if value < 0:
        raise ValueError("worker-gpuid must be a nonnegative GPU index or cpu")
    return value


class RequestHandler:
    """Executes requests on one device (``None``: no GPU).  ``handle`` returns
    the reply of a synchronous command; ``evaluate`` requests go to the
    serving loop (``pump``) and are answered from ``completed``."""

    GEOMETRY_CACHE = 2

    def __init__(self, device, search_parser=None, run_search=None):
        self.device = device
        self._parser = search_parser
        self._run_search = run_search
        self.geometry = OrderedDict()
        self.batch_engine = None
        self.server = None
        self.clock_t0 = time.perf_counter()
        self.requests_served = 0

    # -- helpers ---------------------------------------------------------------
    def _require_gpu(self, what):
        if self.device is None:
            raise RuntimeError(f"{what}: gisapr_amd runs on HIP devices only and this worker has no GPU")

    def _search_args(self, argv):
        from . import search
        parser = (self._parser or search.create_search_parser)()
        args = parser.parse_args(list(argv))
        args.gpuid = self.device.index if self.device is not None else "cpu"
        return args

    def _run(self, args, **kw):
        from . import search
        return (self._run_search or search.run_search)(args, **kw)

    def now(self) -> float:
        return time.perf_counter() - self.clock_t0

    # -- synchronous commands ---------------------------------------------------
    def geometry_request(self, request):
        from .restraint import GeometryRestraint
        self._require_gpu("geometry")
        kwargs = dict(request["geometry_kwargs"], device=str(self.device))
        key = geometry_key(kwargs)
        g = self.geometry.get(key)
        if g is None:
            g = GeometryRestraint(**kwargs)
            self.geometry[key] = g
            while len(self.geometry) > self.GEOMETRY_CACHE:  # bounded over long scans
                self.geometry.popitem(last=False)
        else:
            self.geometry.move_to_end(key)
        overlap, distance = g.evaluate(request["pose"], pivot=request.get("pivot"))
        return int(overlap), float(distance)

    def search_request(self, request, with_model: bool):
        import torch
        from .density import generate_volume
        args = self._search_args(request["argv"])  # argument errors first, as the search program reports them
        self._require_gpu(request["command"])
        kw = dict(calibrate=bool(request.get("calibrate", False)), orientation_rows=request.get("orientation_rows"))
        density_s = 0.0
        volume = None
        if with_model:
            model = dict(request["model_kwargs"], device=str(self.device))
            t0 = time.perf_counter()
            volume = generate_volume(**model)
            torch.cuda.synchronize(self.device)
            density_s = time.perf_counter() - t0
            kw["model_volume"] = volume
        try:
            result = self._run(args, **kw)
        finally:
            kw.clear()  # the volume lives for this request only
            del volume
        if isinstance(result, dict):
            result["density_s"] = density_s
        return result

    def batch_request(self, request):
        import torch
        from .batch import BatchSearchEngine
        args = self._search_args(request["argv"])
        self._require_gpu("search_batch_from_pdb")
        options = request.get("options") or {}
        if self.batch_engine is None or self.batch_engine.options != options:
            self.batch_engine = None
            self.batch_engine = BatchSearchEngine(None, self.device, options)
        try:
            return self.batch_engine.run(args, request["models"], request.get("orientation_rows"),
                                         request.get("calibrate", False))
        except torch.cuda.OutOfMemoryError as exc:
            # drop what the engine holds; a partial batch is never returned
            self.batch_engine = None
            torch.cuda.empty_cache()
            raise RuntimeError("batch evaluation ran out of device memory: lower batch_candidates or "
                               "fixed_density_cache_mb") from exc

    def add_dataset(self, request):
        from .batch import BatchSearchEngine, RigidBody
        from .serve import SearchServer
        self._require_gpu("add_dataset")
        args = self._search_args(request["argv"])
        options = dict(request.get("options") or {})
        path = options.get("path", "split")
        slots = int(options.get("slots", 16))
        if self.server is None or not self.server.datasets:  # settings are free until a dataset is resident
            budget = options.get("hbm_budget_mb")
            self.server = SearchServer(self.device, None if budget is None else int(float(budget) * 2 ** 20),
                                       path=path, slots=slots)
        elif (self.server.path, self.server.slots) != (path, slots):
            raise ValueError(f"this worker serves path={self.server.path!r}, slots={self.server.slots}")
        name = str(request["name"])
        model = request["model_kwargs"]
        body = RigidBody.from_pdb(model["source_pdb"], model["chain_id"], bool(model.get("include_hetatm", False)))
        ctx = BatchSearchEngine(None, self.device, dict(path=path))._context(args)
        ds = self.server.add_dataset(name, ctx["engine"], ctx["ps"], body, float(model["res"]))
        return dict(name=name, resident_bytes=int(ds.resident_bytes), particles=int(ctx["ps"].B),
                    directions=len(ctx["engine"].dir_rot), used_bytes=int(self.server.used),
                    budget_bytes=int(self.server.budget))

    def ping(self):
        import torch
        out = dict(pid=os.getpid(), device=str(self.device) if self.device is not None else "cpu",
                   geometry_cache_entries=len(self.geometry), requests_served=self.requests_served)
        if self.batch_engine is not None:
            out.update(batch_particle_cache_bytes=self.batch_engine.particles.bytes,
                       batch_volume_cache_bytes=self.batch_engine.volumes.bytes)
        if self.server is not None:
            out.update(datasets=sorted(self.server.datasets), served_bytes=int(self.server.used),
                       served_budget_bytes=int(self.server.budget), served_steps=int(self.server.steps))
        if self.device is not None:
            out.update(device_peak_allocated_bytes=int(torch.cuda.max_memory_allocated(self.device)),
                       device_peak_reserved_bytes=int(torch.cuda.max_memory_reserved(self.device)))
        return out

    def handle(self, request, initial_cwd):
        """(reply or None, stop).  ``evaluate`` replies later (None here)."""
        command = None
        try:
            if not isinstance(request, dict):
                raise TypeError("worker request must be a dictionary")
            command = request.get("command")
            os.chdir(request.get("cwd") or initial_cwd)
            if command == "shutdown":
                return {"ok": True, "result": None}, True
            if command == "evaluate":
                self.submit(request)
                return None, False
            import torch
            with torch.inference_mode():
                if command == "geometry":
                    result = self.geometry_request(request)
                elif command == "generate_mrc":
                    from .density import generate_mrc
                    self._require_gpu("generate_mrc")
                    result = generate_mrc(**dict(request["mrc_kwargs"], device=str(self.device)))
                elif command in ("search", "search_from_pdb"):
                    result = self.search_request(request, command == "search_from_pdb")
                elif command == "search_batch_from_pdb":
                    result = self.batch_request(request)
                elif command == "add_dataset":
                    result = self.add_dataset(request)
                elif command == "ping":
                    result = self.ping()
                else:
                    raise ValueError(f"Unknown worker command: {command!r}")
            self.requests_served += 1
            return {"ok": True, "result": result}, False
        except (Exception, SystemExit) as error:  # argparse exits on bad job arguments
            reply = {"ok": False, "error": f"{type(error).__name__}: {error}", "traceback": traceback.format_exc()}
            if command == "evaluate" and isinstance(request, dict):
                reply["rid"] = request.get("rid")
            return reply, False

    # -- serving ---------------------------------------------------------------
    def submit(self, request):
        from .serve import Request
        self._require_gpu("evaluate")
        if self.server is None:
            raise KeyError(f"dataset {request.get('dataset')!r} is not resident (no add_dataset yet)")
        self.server.submit(Request(request["rid"], str(request["dataset"]), request["poses"], self.now()))

    def busy(self) -> bool:
        return self.server is not None and self.server.busy()

    def pump(self):
        """One pass of the serving loop; returns the replies of the requests
        completed by it."""
        if self.server is None:
            return False, []
        import torch
        with torch.inference_mode():
            launched = self.server.pump(self.now)
        replies = []
        for req in self.server.take_completed():
            self.requests_served += 1
            replies.append({"ok": True, "rid": req.rid,
                            "result": {"scores": [float(v) for v in req.scores], "latency_s": float(req.latency)}})
        return launched, replies


def run_worker(create_search_parser=None, run_search=None, choose_device=None, argv=None) -> None:
    """Connect to the controller, then execute its requests until shutdown or
    disconnect (the reference's run_worker signature: the search program's
    parser, run_search and device chooser, each defaulting to this
    package's)."""
    from multiprocessing.connection import Client
    args = create_worker_parser().parse_args(argv)
    host, port = args.worker_connect
    if host not in LOOPBACK:
        raise ValueError("the worker connects to a loopback host only")
    authkey = bytes.fromhex(args.worker_auth)
    if not authkey:
        raise ValueError("worker authentication key cannot be empty")
    gpu = _parse_gpu(args.worker_gpuid)
    device = None
    if gpu != "cpu":
        if choose_device is None:
            from .search import choose_device
        device = choose_device(gpu)
        from ._lib import check, lib
        check(lib().gis_init(device.index if device.index is not None else 0), "gis_init")
    handler = RequestHandler(device, create_search_parser, run_search)
    initial_cwd = os.getcwd()
    with Client((host, int(port)), family="AF_INET", authkey=authkey) as conn:
        conn.send({"ok": True, "event": "ready", "pid": os.getpid(), "device": str(device or "cpu")})
        stop = False
        while not stop:
            try:
                # wait for messages only while no step is in flight
                if conn.poll(0.0002 if handler.busy() else 0.05):
                    request = conn.recv()
                    reply, stop = handler.handle(request, initial_cwd)
                    if stop:
                        reply_stop, reply = reply, None
                    if reply is not None:
                        sys.stdout.flush()
                        sys.stderr.flush()
                        conn.send(reply)
                if stop:  # finish the requests already accepted, then confirm
                    replies = []
                    while handler.busy():
                        launched, done = handler.pump()
                        replies += done
                        if not launched:
                            time.sleep(0.0002)
                    for r in replies:
                        conn.send(r)
                    conn.send(reply_stop)
                    break
                _, replies = handler.pump()
                for reply in replies:
                    conn.send(reply)
            except (EOFError, ConnectionError, BrokenPipeError):
                break


def main(argv=None) -> None:
    run_worker(argv=argv)


if __name__ == "__main__":
    main()
