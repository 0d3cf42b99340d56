"""Request server (qm_server_*) on the GPU: every result equals the same
request run alone, the oracle agrees (energies within 1e-10 * sum |c|, samples
identical for the same uniforms), and admission keeps the footprint within the
budget."""
from __future__ import annotations

import threading

import numpy as np
import pytest

from qulacs_amd.serving import Server
from qulacs_amd.workloads import serving_requests
from tests.batch_util import oracle_run

pytestmark = pytest.mark.gpu


def _requests():
    rs = serving_requests(14, rate=1000.0, seed=11, n_range=(3, 22), depth_range=(2, 8), samples=300)
    # one of each path: tiny (padded to 4 qubits), batched, fused large
    return rs


def _run_all(srv, rs):
    ids = [srv.submit(r.n, r.gates, r.terms, r.uniforms) for r in rs]
    return [srv.wait(i) for i in ids]


def test_results_equal_alone_and_in_the_mix(oracle):
    rs = _requests()
    print("sizes", [r.n for r in rs], flush=True)
    with Server(hbm_budget_bytes=1 << 33, large_slots=3, small_max_qubits=16) as srv:
        mixed = _run_all(srv, rs)
        st = srv.stats()
        print("stats", st, flush=True)
        assert st["completed"] == len(rs) and st["failed"] == 0
    for r, (e, smp, lat) in zip(rs, mixed):
        with Server(hbm_budget_bytes=1 << 33, large_slots=1, small_max_qubits=16) as alone_srv:
            e1, s1, _ = _run_all(alone_srv, [r])[0]
        assert e == e1 and np.array_equal(smp, s1), r.n
        psi, want = oracle_run(oracle, r.n, r.gates, r.terms)
        scale = max(1.0, sum(abs(t[3]) for t in r.terms))
        assert abs(e - want) < 1e-10 * scale, (r.n, e, want)
        assert np.array_equal(smp, oracle.sampling(psi, r.uniforms)), r.n
        assert lat > 0


def test_admission_respects_budget_and_threads_submit():
    rs = serving_requests(24, rate=1000.0, seed=12, n_range=(14, 22), depth_range=(2, 6), samples=64)
    budget = (16 << 22) * 2  # room for two 22-qubit states
    out = {}
    with Server(hbm_budget_bytes=budget, large_slots=4, small_max_qubits=15) as srv:
        def client(chunk):
            for k, r in chunk:
                out[k] = srv.wait(srv.submit(r.n, r.gates, r.terms, r.uniforms))
        th = [threading.Thread(target=client, args=(list(enumerate(rs))[i::3],)) for i in range(3)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        st = srv.stats()
    print("stats", st, flush=True)
    assert len(out) == len(rs) and st["failed"] == 0
    assert st["peak_bytes"] <= budget


def test_invalid_requests_rejected():
    from qulacs_amd._lib import QulacsMI355Error

    with Server(hbm_budget_bytes=1 << 24, large_slots=1) as srv:
        with pytest.raises(QulacsMI355Error):
            srv.submit(21, [], [], [0.5])  # 32 MiB state over a 16 MiB budget
        with pytest.raises(QulacsMI355Error):
            srv.submit(5, [("dense", [7], np.eye(2))], [], [0.5])
        e, s, _ = srv.wait(srv.submit(5, [], [(0, 1, 0, 2.0)], [0.25, 0.75]))
        assert e == 2.0 and list(s) == [0, 0]
