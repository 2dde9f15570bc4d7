"""Serving loop (pwmhip_server_*): every request's result -- all level words,
zeros and query answers -- is the same when the request runs alone and when
it runs mixed with others (continuous batching, several streams, queries
batched across indices); HBM-budget admission holds builds back until
memory frees up."""
import threading
import time

import numpy as np
import pytest

import oracle_lib as O

pwm_amd = pytest.importorskip("pwm_amd")

ALPHABETS = [(5, 3), (20, 5), (256, 8)]


def _requests(seed, count, max_len=1 << 21):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        sigma, L = ALPHABETS[rng.choice(3, p=[0.6, 0.3, 0.1])]
        n = int(np.clip(rng.lognormal(np.log(64 << 10), 1.5), 1 << 10, max_len))
        if i % 13 == 5:
            n = int(rng.integers(1, 100))  # tiny texts mixed in
        text = rng.integers(0, sigma, n).astype(np.uint8)
        algo = ["ppc", "pps", "dd_pc", "pc"][i % 4]
        tree = bool(i % 3 == 1)
        q = 2000
        op = rng.integers(0, 3, q).astype(np.uint32)
        a = np.where(op == 0, rng.integers(0, n, q), text[rng.integers(0, n, q)]).astype(np.uint64)
        b = np.where(op == 1, rng.integers(0, n + 1, q), rng.integers(1, 50, q)).astype(np.uint64)
        reqs.append(dict(text=text, L=L, algo=algo, tree=tree, op=op, a=a, b=b))
    return reqs


def _run(srv, reqs, order, stagger=None):
    """Submits builds and their queries in `order`; returns per-request
    (levels, zeros, answers)."""
    ids = {}

    def submit():
        for i in order:
            r = reqs[i]
            bid = srv.build(r["text"], r["L"], r["algo"], r["tree"])
            qid = srv.query(bid, r["op"], r["a"], r["b"])  # queued until the index is ready
            ids[i] = (bid, qid)
            if stagger is not None:
                time.sleep(stagger[i])

    if stagger is None:
        submit()
    else:
        t = threading.Thread(target=submit)
        t.start()
        t.join()
    out = {}
    for i in order:
        bid, qid = ids[i]
        srv.wait(qid)
        srv.wait(bid)
        out[i] = (srv.levels(bid), srv.zeros(bid), srv.result(qid))
        srv.release(bid)
    return out


def _alone(reqs):
    srv = pwm_amd.Server(workers=1, max_batch_bytes=64 << 20, max_query_batch=1 << 20)
    out = {}
    try:
        for i in range(len(reqs)):
            out.update(_run(srv, reqs, [i]))
    finally:
        srv.close()
    return out


@pytest.mark.gpu
def test_serve_alone_equals_mixed():
    reqs = _requests(77, 48)
    alone = _alone(reqs)
    # results alone equal a plain build and brute-force answers
    for i in range(0, len(reqs), 7):
        r = reqs[i]
        ws = pwm_amd.build(r["text"], r["L"], r["algo"], r["tree"])
        for x, y in zip(alone[i][0], ws.bvs()):
            assert np.array_equal(x, y), i
        assert np.array_equal(alone[i][1], ws.zeros()), i
        ws.free()
        want = O.symbol_queries(r["text"], r["op"], r["a"], r["b"])
        assert np.array_equal(alone[i][2], want), i

    srv = pwm_amd.Server(workers=3, max_batch_bytes=64 << 20, max_query_batch=1 << 20)
    try:
        rng = np.random.default_rng(5)
        order = list(rng.permutation(len(reqs)))
        # a burst, then Poisson-like arrivals while earlier work is on the GPU
        mixed = _run(srv, reqs, order[:24])
        stagger = {i: float(rng.exponential(0.002)) for i in order[24:]}
        mixed.update(_run(srv, reqs, order[24:], stagger))
        st = srv.stats()
    finally:
        srv.close()
    for i in range(len(reqs)):
        la, za, qa = alone[i]
        lm, zm, qm = mixed[i]
        assert len(la) == len(lm), i
        for x, y in zip(la, lm):
            assert np.array_equal(x, y), i
        assert np.array_equal(za, zm), i
        assert np.array_equal(qa, qm), i
    assert st["builds_done"] == len(reqs)
    assert st["build_launches"] < len(reqs)  # builds were batched
    assert st["mean_texts_per_launch"] > 1.0
    assert st["query_requests_done"] == len(reqs)


@pytest.mark.gpu
def test_serve_budget_admission_and_errors():
    rng = np.random.default_rng(3)
    small = dict(workers=1, max_batch_texts=4, max_batch_bytes=4 << 20, max_query_batch=1 << 16)
    probe = pwm_amd.Server(**small, hbm_budget_bytes=1 << 40)
    base = probe.stats()["bytes_in_use"]  # the worker's scratch allowance
    probe.close()
    n = 8 << 20  # 8 MiB, 8 levels: about 8.6 MB resident
    srv = pwm_amd.Server(**small, hbm_budget_bytes=base + 12_000_000)
    try:
        t1 = rng.integers(0, 256, n).astype(np.uint8)
        t2 = rng.integers(0, 256, n).astype(np.uint8)
        a = srv.build(t1, 8)
        srv.wait(a)
        b = srv.build(t2, 8)  # does not fit while a is resident
        with pytest.raises(pwm_amd.PwmError) as e:
            srv.wait(b, timeout_ms=500)
        assert e.value.status == pwm_amd.ERR_TIMEOUT
        srv.release(a)
        srv.wait(b, timeout_ms=60000)
        ws = pwm_amd.build(t2, 8, "ppc")
        for x, y in zip(srv.levels(b), ws.bvs()):
            assert np.array_equal(x, y)
        ws.free()
        # larger than the whole budget: refused at once
        with pytest.raises(pwm_amd.PwmError) as e:
            srv.build(np.zeros(64 << 20, np.uint8), 8)
        assert e.value.status == pwm_amd.ERR_BUDGET
        # symbol out of range: the request fails, the server goes on
        bad = srv.build(np.array([1, 2, 200], np.uint8), 3, check_range=True)
        with pytest.raises(pwm_amd.PwmError) as e:
            srv.wait(bad)
        assert e.value.status == pwm_amd.ERR_SYMBOL_RANGE
        ok = srv.build(np.array([1, 2, 7], np.uint8), 3)
        srv.wait(ok)
        assert srv.stats()["bytes_peak"] <= base + 12_000_000
    finally:
        srv.close()
