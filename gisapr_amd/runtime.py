"""Controller side of the worker protocol: start one persistent worker process
per GPU and send it requests (gisapr_runtime.py:153-278's PermanentWorker,
get_worker and shutdown_workers, with the same request/reply semantics).

The controller listens on an ephemeral loopback port with a fresh random
key, starts ``python -u -m gisapr_amd.worker --worker-connect 127.0.0.1
<port> --worker-auth <hex> --worker-gpuid <gpu>`` and waits for its "ready"
greeting.  ``request`` sends one command and waits for its reply, raising
RuntimeError with the worker's message and traceback on failure, or when
the process exits meanwhile.  For serving, ``submit`` sends an ``evaluate``
request without waiting and ``collect`` gathers the replies, which arrive
as their requests complete (in any order).
"""
from __future__ import annotations

import atexit
import itertools
import os
import secrets
import subprocess
import sys
import threading
import time
from multiprocessing.connection import Listener
from typing import Dict, List, Optional


class WorkerError(RuntimeError):
    pass


class PermanentWorker:
    """One worker process bound to one GPU (``gpuid`` an index, or "cpu")."""

    def __init__(self, gpuid=0, startup_timeout: float = 900.0, worker_id: int = 0, module: str = "gisapr_amd.worker",
                 env: Optional[dict] = None):
        self.lock = threading.RLock()
        self.process = None
        self.connection = None
        self.gpuid = str(gpuid)
        self.worker_id = int(worker_id)
        self.slot = (self.gpuid, self.worker_id)
        self._rids = itertools.count()
        self._early: Dict[object, dict] = {}   # evaluate replies received while waiting for another reply
        self.outstanding = set()
        key = secrets.token_bytes(32)
        listener = Listener(("127.0.0.1", 0), family="AF_INET", authkey=key)
        accepted, failed = [], []

        def accept():
            try:
                accepted.append(listener.accept())
            except Exception as exc:  # reported by the waiting loop below
                failed.append(exc)

        thread = threading.Thread(target=accept, daemon=True)
        thread.start()
        host, port = listener.address
        cmd = [sys.executable, "-u", "-m", module, "--worker-connect", host, str(port), "--worker-auth", key.hex(),
               "--worker-gpuid", self.gpuid]
        child_env = dict(os.environ if env is None else env)
        root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
        child_env["PYTHONPATH"] = root + (os.pathsep + child_env["PYTHONPATH"] if child_env.get("PYTHONPATH") else "")
        try:
            self.process = subprocess.Popen(cmd, env=child_env)
            deadline = time.monotonic() + float(startup_timeout)
            while not accepted:
                if failed:
                    raise WorkerError(f"worker connection failed: {failed[0]}")
                if self.process.poll() is not None:
                    raise WorkerError(f"worker {self.slot} exited with code {self.process.returncode}")
                if time.monotonic() >= deadline:
                    raise TimeoutError(f"worker {self.slot} did not connect within {startup_timeout}s")
                thread.join(0.1)
            self.connection = accepted[0]
            while not self.connection.poll(0.1):
                if self.process.poll() is not None or time.monotonic() >= deadline:
                    raise WorkerError(f"worker {self.slot} did not become ready")
            ready = self.connection.recv()
            if not (isinstance(ready, dict) and ready.get("ok") and ready.get("event") == "ready"):
                raise WorkerError(f"invalid worker greeting: {ready!r}")
            self.pid = ready.get("pid")
            self.device = ready.get("device")
        except BaseException:
            self.close()
            raise
        finally:
            listener.close()

    def _recv(self, what: str) -> dict:
        while not self.connection.poll(0.5):
            if self.process.poll() is not None:
                raise WorkerError(f"worker {self.slot} died during {what}")
        try:
            return self.connection.recv()
        except EOFError as exc:
            raise WorkerError(f"worker {self.slot} disconnected during {what}") from exc

    @staticmethod
    def _unwrap(reply: dict):
        if not reply.get("ok"):
            raise WorkerError(f"{reply.get('error')}\n{reply.get('traceback', '')}")
        return reply.get("result")

    def request(self, command: str, **payload):
        """Send one synchronous command; its result, or WorkerError."""
        with self.lock:
            if self.process.poll() is not None:
                raise WorkerError(f"worker {self.slot} exited with code {self.process.returncode}")
            self.connection.send({"command": command, "cwd": os.getcwd(), **payload})
            while True:
                reply = self._recv(command)
                if "rid" in reply:  # an evaluate reply that completed meanwhile
                    self._early[reply["rid"]] = reply
                    continue
                return self._unwrap(reply)

    def submit(self, dataset: str, poses, rid=None):
        """Queue an ``evaluate`` request; returns its id."""
        import numpy as np
        rid = next(self._rids) if rid is None else rid
        with self.lock:
            self.connection.send({"command": "evaluate", "rid": rid, "dataset": dataset,
                                  "poses": np.asarray(poses, np.float64).reshape(-1, 6).tolist()})
            self.outstanding.add(rid)
        return rid

    def collect(self, count: Optional[int] = None, timeout: Optional[float] = None) -> List[dict]:
        """Replies of submitted requests ({"rid", "ok", "result" or "error"}):
        at least ``count`` of them (all outstanding when None), waiting at
        most ``timeout`` seconds."""
        want = len(self.outstanding) if count is None else int(count)
        deadline = None if timeout is None else time.monotonic() + float(timeout)
        out = []
        with self.lock:
            for rid in list(self._early):
                if rid in self.outstanding:
                    out.append(self._early.pop(rid))
                    self.outstanding.discard(rid)
            while len(out) < want:
                if deadline is not None and time.monotonic() >= deadline:
                    raise TimeoutError(f"{want - len(out)} evaluate replies still pending")
                if not self.connection.poll(0.05):
                    if self.process.poll() is not None:
                        raise WorkerError(f"worker {self.slot} died while serving")
                    continue
                reply = self.connection.recv()
                if "rid" not in reply:
                    raise WorkerError(f"unexpected reply while collecting: {reply!r}")
                self.outstanding.discard(reply["rid"])
                out.append(reply)
        return out

    def close(self):
        """Ask the worker to shut down (it finishes accepted evaluations
        first), then end the process if it does not exit."""
        if self.connection is not None:
            try:
                if self.process is not None and self.process.poll() is None:
                    self.connection.send({"command": "shutdown"})
            except (OSError, EOFError):
                pass
        if self.process is not None:
            try:
                self.process.wait(timeout=30)
            except subprocess.TimeoutExpired:
                self.process.terminate()
                try:
                    self.process.wait(timeout=5)
                except subprocess.TimeoutExpired:
                    self.process.kill()
                    self.process.wait()
        if self.connection is not None:
            self.connection.close()
            self.connection = None


_workers: Dict[tuple, PermanentWorker] = {}
_registry_lock = threading.Lock()


def get_worker(gpuid=0, worker_id: int = 0, startup_timeout: float = 900.0) -> PermanentWorker:
    """The process-wide worker of (gpuid, worker_id), started on first use."""
    key = (str(gpuid), int(worker_id))
    with _registry_lock:
        if key not in _workers:
            _workers[key] = PermanentWorker(gpuid, startup_timeout, worker_id)
        return _workers[key]


def shutdown_workers() -> None:
    with _registry_lock:
        workers = list(_workers.values())
        _workers.clear()
    for w in workers:
        w.close()


atexit.register(shutdown_workers)
