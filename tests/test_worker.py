"""The worker protocol on the CPU (gisapr_worker.py:39-173,
gisapr_runtime.py:153-278): a worker process without a GPU answers over its
authenticated loopback connection, reports failed requests without dying,
refuses compute commands (HIP devices only), and shuts down on request."""
import os
import time

import pytest

from gisapr_amd import worker
from gisapr_amd.runtime import PermanentWorker, WorkerError


@pytest.fixture(scope="module")
def cpu_worker():
    w = PermanentWorker("cpu", startup_timeout=300)
    yield w
    w.close()


def test_ready_and_ping(cpu_worker):
    info = cpu_worker.request("ping")
    assert info["pid"] == cpu_worker.pid == cpu_worker.process.pid
    assert info["device"] == "cpu" and info["geometry_cache_entries"] == 0


def test_failures_keep_the_worker(cpu_worker):
    with pytest.raises(WorkerError, match="Unknown worker command"):
        cpu_worker.request("no_such_command")
    # argparse's exit on bad job arguments is reported, not fatal
    with pytest.raises(WorkerError, match="SystemExit"):
        cpu_worker.request("search", argv=["--no-such-flag"])
    # well-formed compute requests fail loudly without a GPU
    argv = ["--i", "x", "--p", "p.star", "--o", "o.txt", "--ang", "a.star", "--mrc", "m.mrc"]
    with pytest.raises(WorkerError, match="HIP devices only"):
        cpu_worker.request("search_from_pdb", argv=argv, model_kwargs=dict(pdb_path="m.pdb", boxsize=32, apix=1.0,
                                                                          res=2.0))
    with pytest.raises(WorkerError, match="HIP devices only"):
        cpu_worker.request("geometry", geometry_kwargs=dict(pdb_path=__file__, chain_id="A"), pose=[0] * 6)
    with pytest.raises(WorkerError, match="TypeError"):
        cpu_worker.connection.send(["not", "a", "dict"])
        cpu_worker._unwrap(cpu_worker._recv("bad request"))
    assert cpu_worker.request("ping")["requests_served"] >= 1


def test_evaluate_errors_carry_the_request_id(cpu_worker):
    rid = cpu_worker.submit("nowhere", [[0, 0, 0, 0, 0, 0]])
    (reply,) = cpu_worker.collect(timeout=60)
    assert reply["rid"] == rid and not reply["ok"] and "HIP devices only" in reply["error"]
    assert not cpu_worker.outstanding


def test_cwd_per_request(cpu_worker, tmp_path):
    # the job's cwd is applied; the next job without one starts in the worker's own
    cpu_worker.connection.send({"command": "ping", "cwd": str(tmp_path)})
    assert cpu_worker._unwrap(cpu_worker._recv("ping"))["pid"] == cpu_worker.pid
    with pytest.raises(WorkerError, match="FileNotFoundError"):
        cpu_worker.connection.send({"command": "ping", "cwd": str(tmp_path / "missing")})
        cpu_worker._unwrap(cpu_worker._recv("ping"))


def test_shutdown_ends_the_process():
    w = PermanentWorker("cpu", startup_timeout=300)
    assert w.request("ping")["device"] == "cpu"
    assert w.request("shutdown") is None
    assert w.process.wait(timeout=60) == 0
    w.close()


def test_connection_rules():
    with pytest.raises(ValueError, match="loopback"):
        worker.run_worker(argv=["--worker-connect", "10.0.0.1", "5000", "--worker-auth", "ab"])
    with pytest.raises(ValueError, match="empty"):
        worker.run_worker(argv=["--worker-connect", "127.0.0.1", "5000", "--worker-auth", ""])
    with pytest.raises(ValueError, match="nonnegative"):
        worker.run_worker(argv=["--worker-connect", "127.0.0.1", "5000", "--worker-auth", "ab",
                                "--worker-gpuid", "-1"])


def test_geometry_key_follows_the_file(tmp_path):
    p = tmp_path / "m.pdb"
    p.write_text("ATOM\n")
    k1 = worker.geometry_key(dict(pdb_path=str(p), chain_id="A", boxsize=64))
    assert worker.geometry_key(dict(pdb_path=str(tmp_path / "." / "m.pdb"), chain_id="A", boxsize=64)) == k1
    assert worker.geometry_key(dict(pdb_path=str(p), chain_id="B", boxsize=64)) != k1
    time.sleep(0.01)
    p.write_text("ATOM\nATOM\n")
    os.utime(p)
    assert worker.geometry_key(dict(pdb_path=str(p), chain_id="A", boxsize=64)) != k1
