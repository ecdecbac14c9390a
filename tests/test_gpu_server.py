"""GPU tests of the serving loop (phem_server_*): every request's result in a
mixed, concurrently batched stream equals the result of the same request
served alone and the oracle's Kruskal forest, labels and exact weight; a
request larger than a slot is refused; statistics add up."""
import os
import threading

import numpy as np
import pytest

import oracle as O
import synth as S

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def P():
    import phem_native
    phem_native.require_device()
    return phem_native


def host_graph(g, seed=3):
    """Small graphs of the serving mix's shapes (sizes scaled down)."""
    rng = np.random.default_rng(seed * 1000 + g)
    kind = g % 3
    if kind == 0:
        sc = int(rng.integers(3, 11))
        u, v, w = S.rmat(sc, int(rng.integers(1, 16)) << (sc - 1), S.stream_seed(seed, g))
        return 1 << sc, u, v, w
    if kind == 1:
        n = int(rng.integers(1, 2000))
        u, v, w = S.uniform(n, int(rng.integers(0, 8 * n)), S.stream_seed(seed, g))
        return n, u, v, w
    side = int(rng.integers(2, 30))
    u, v, w = S.grid(side, S.stream_seed(seed, g))
    return side * side, u, v, w


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int32)).cuda()


def submit(srv, n, u, v, w, on_device, arrival=-1.0):
    ids = np.zeros(max(n, 1), np.uint32)
    lab = np.zeros(max(n, 1), np.uint32)
    if on_device:
        t = srv.submit(n, dev(u), dev(v), dev(w), ids, lab, arrival)
    else:
        t = srv.submit(n, u, v, w, ids, lab, arrival)
    return t, ids, lab


def test_mix_equals_alone_and_oracle(P):
    graphs = [host_graph(g) for g in range(120)]
    # served alone: one request at a time through a one-graph server
    alone = []
    solo = P.Server(mem_bytes=256 << 20, slots=1, max_graphs=1)
    for n, u, v, w in graphs:
        t, ids, lab = submit(solo, n, u, v, w, on_device=False)
        r = solo.wait(t)
        assert r["status"] == 0 and r["batch_graphs"] == 1
        alone.append((ids[: r["forest_count"]].copy(), lab[:n].copy(), r))
    solo.close()
    # the mix: all submitted at once, alternating host and device inputs, from
    # two submitting threads, several batches in flight
    srv = P.Server(mem_bytes=512 << 20, slots=3, max_graphs=16, window_us=200)
    out = [None] * len(graphs)

    def feed(idx):
        for g in idx:
            n, u, v, w = graphs[g]
            out[g] = submit(srv, n, u, v, w, on_device=(g % 2 == 0))

    th = [threading.Thread(target=feed, args=(range(k, len(graphs), 2),)) for k in (0, 1)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    batched = 0
    for g, (n, u, v, w) in enumerate(graphs):
        t, ids, lab = out[g]
        r = srv.wait(t)
        assert r["status"] == 0
        batched += r["batch_graphs"] > 1
        got_ids = ids[: r["forest_count"]]
        a_ids, a_lab, a_r = alone[g]
        assert np.array_equal(got_ids, a_ids), f"graph {g}"
        assert np.array_equal(lab[:n], a_lab), f"graph {g}"
        assert (r["weight_fx"], r["weight_exact"]) == (a_r["weight_fx"], a_r["weight_exact"])
        k_ids, k_lab = O.kruskal(n, u, v, w, 1)
        assert np.array_equal(got_ids, k_ids), f"graph {g}"
        assert np.array_equal(lab[:n], k_lab), f"graph {g}"
        ws, exact = O.weight_sum(k_ids, w, 1)
        assert bool(r["weight_exact"]) == exact
        if exact:
            assert r["weight_fx"] == ws
        assert r["arrival"] <= r["start"] <= r["done"]
    assert batched > 0, "no request shared a batch"
    st = srv.stats()
    assert st["completed"] == st["submitted"] == len(graphs)
    assert st["edges"] == sum(len(gr[1]) for gr in graphs)
    assert 0 < st["lat_p50"] <= st["lat_p99"] <= st["lat_max"]
    assert st["mean_batch_graphs"] > 1
    srv.close()


def test_oversized_request_refused(P):
    srv = P.Server(mem_bytes=64 << 20, slots=1, max_graphs=8)
    n, m = 1 << 16, 1 << 22
    u, v, w = S.uniform(n, m, 1)
    with pytest.raises(P.PhemError) as e:
        srv.submit(n, u, v, w)
    assert e.value.code == P.EBUSY
    assert srv.stats()["refused"] == 1
    # a small one still goes through
    n2, u2, v2, w2 = host_graph(1)
    t, ids, lab = submit(srv, n2, u2, v2, w2, on_device=False)
    r = srv.wait(t)
    k_ids, _ = O.kruskal(n2, u2, v2, w2, 1)
    assert np.array_equal(ids[: r["forest_count"]], k_ids)
    srv.close()


def test_failed_allocation_is_reported_and_forgotten(P):
    """A server whose slot memory cannot be allocated (256 TiB) is refused
    with PHEM_ENOMEM, and the failed allocation does not leak into the next
    call's launch checks: a solve right after it succeeds."""
    with pytest.raises(P.PhemError) as e:
        P.Server(mem_bytes=1 << 48, slots=1)
    assert e.value.code == P.ENOMEM
    n, u, v, w = host_graph(2)
    ids, lab, st = P.mst_solve(n, torch.from_numpy(u.view(np.int32)).cuda(),
                               torch.from_numpy(v.view(np.int32)).cuda(),
                               torch.from_numpy(w.view(np.int32)).cuda(), P.W_F32)
    torch.cuda.synchronize()
    k_ids, _ = O.kruskal(n, u, v, w, 1)
    assert np.array_equal(ids.cpu().numpy().view(np.uint32), k_ids)


def test_admission_by_footprint_splits_batches(P):
    """With room for only a few graphs per slot, a burst is served in many
    batches, each within the slot's memory, all correct."""
    graphs = [host_graph(g, seed=5) for g in range(40)]
    srv = P.Server(mem_bytes=36 << 20, slots=2, max_graphs=4096)
    out = [submit(srv, n, u, v, w, on_device=False) for n, u, v, w in graphs]
    srv.drain()
    for g, (n, u, v, w) in enumerate(graphs):
        t, ids, lab = out[g]
        r = srv.poll(t)
        assert r is not None and r["status"] == 0
        k_ids, k_lab = O.kruskal(n, u, v, w, 1)
        assert np.array_equal(ids[: r["forest_count"]], k_ids)
        assert np.array_equal(lab[:n], k_lab)
    st = srv.stats()
    assert st["batches"] > 2
    srv.close()


def test_unknown_ticket(P):
    srv = P.Server(mem_bytes=16 << 20, slots=1)
    with pytest.raises(P.PhemError) as e:
        srv.wait(12345)
    assert e.value.code == P.ENOTFOUND
    srv.close()


def test_u32_weights_server(P):
    srv = P.Server(mem_bytes=64 << 20, slots=2, max_graphs=64, wmode=P.W_U32)
    rng = np.random.default_rng(2)
    reqs = []
    for g in range(20):
        n = int(rng.integers(2, 500))
        m = int(rng.integers(0, 3000))
        u = rng.integers(0, n, m).astype(np.uint32)
        v = rng.integers(0, n, m).astype(np.uint32)
        w = rng.integers(0, 8, m).astype(np.uint32)  # heavy ties
        reqs.append((n, u, v, w) + submit(srv, n, u, v, w, on_device=False))
    for n, u, v, w, t, ids, lab in reqs:
        r = srv.wait(t)
        k_ids, k_lab = O.kruskal(n, u, v, w, 0)
        assert np.array_equal(ids[: r["forest_count"]], k_ids)
        assert np.array_equal(lab[:n], k_lab)
        assert r["weight_fx"] == int(w[k_ids].astype(np.uint64).sum())
    srv.close()


@pytest.mark.slow
def test_serving_mix_4096_poisson_vs_oracle(P):
    """The C5 serving mix at its workload size: all 4096 graphs (n
    log-uniform in [2^8, 2^20], degree 4..32, half uniform and half RMAT,
    device resident, 4.7 G edges) submitted at Poisson arrival times through a
    three-slot server, as the benchmark does.  Every result equals the
    oracle: the forest ids equal the PBBS filter-Kruskal restatement's
    (MST/src/utils.cu:794-841, (weight, id) order), the exact weight its
    integer sum, and the labels the CPU union-find's minimum ids."""
    import time

    import phem_workloads as W
    graphs_n, seed, rate = 4096, 5, 8000.0
    t0 = time.perf_counter()
    graphs = [W.mix_graph(seed, g) for g in range(graphs_n)]
    torch.cuda.synchronize()
    edges = sum(g[1].numel() for g in graphs)
    print(f"mix: {graphs_n} graphs, {edges} edges, generated in "
          f"{time.perf_counter() - t0:.1f} s", flush=True)
    srv = P.Server(mem_bytes=0, slots=3, max_graphs=4096, max_batch_edges=1 << 27,
                   wmode=P.W_F32)
    outs = [(np.zeros(g[0], np.uint32), np.zeros(g[0], np.uint32)) for g in graphs]
    arrivals = np.cumsum(np.random.default_rng(seed).exponential(1.0 / rate, graphs_n))
    base = P.Server.clock()
    ticks = []
    for i, (n, u, v, w) in enumerate(graphs):
        at = base + arrivals[i]
        while P.Server.clock() < at:
            pass
        ticks.append(srv.submit(n, u, v, w, outs[i][0], outs[i][1], at))
    res = [srv.wait(t) for t in ticks]
    st = srv.stats()
    srv.close()
    print(f"served in {st['batches']} batches (mean {st['mean_batch_graphs']:.1f} graphs), "
          f"p50 {st['lat_p50'] * 1e3:.2f} ms, p99 {st['lat_p99'] * 1e3:.2f} ms", flush=True)
    assert st["completed"] == graphs_n and st["batches"] > 1
    t0 = time.perf_counter()

    def check(i):
        """Graph i against the oracle, single-threaded (the oracle's C calls
        release the interpreter lock, so graphs are checked in parallel)."""
        (n, u, v, w), r = graphs[i], res[i]
        assert r["status"] == 0, f"graph {i}: {r['error']}"
        hu, hv, hw = (t.cpu().numpy().view(np.uint32) for t in (u, v, w))
        want = O.par_mst(n, hu, hv, hw, 1, 1)
        k = r["forest_count"]
        assert np.array_equal(outs[i][0][:k], want), f"graph {i} ids"
        wfx, exact = O.weight_sum(want, hw, 1)
        assert r["weight_fx"] == wfx and bool(r["weight_exact"]) == exact, f"graph {i} weight"
        assert np.array_equal(outs[i][1], O.par_cc(n, hu, hv, 1)), f"graph {i} labels"

    import concurrent.futures as cf
    workers = max(1, min(16, len(os.sched_getaffinity(0))))
    with cf.ThreadPoolExecutor(workers) as pool:
        for done, _ in enumerate(pool.map(check, range(graphs_n)), 1):
            if done % 512 == 0:
                print(f"checked {done} graphs against the oracle in "
                      f"{time.perf_counter() - t0:.1f} s ({workers} threads)", flush=True)


def test_wait_claims_its_ticket(P):
    """A ticket being waited for is claimed: a poll or second wait meanwhile
    gets PHEM_EBUSY instead of freeing the request under the waiter.  The
    server's batching window (300 ms for a lone request) keeps the request
    pending while the waiter thread, started 50 ms earlier, holds it."""
    import time
    srv = P.Server(mem_bytes=64 << 20, slots=1, max_graphs=4, window_us=300000)
    n, u, v, w = host_graph(0)
    t, ids, lab = submit(srv, n, u, v, w, on_device=False)
    box = {}
    th = threading.Thread(target=lambda: box.setdefault("r", srv.wait(t)))
    th.start()
    time.sleep(0.05)
    for call in (lambda: srv.poll(t), lambda: srv.wait(t)):
        with pytest.raises(P.PhemError) as e:
            call()
        assert e.value.code == P.EBUSY
    th.join()
    assert box["r"]["status"] == 0 and box["r"]["error"] == ""
    k_ids, _ = O.kruskal(n, u, v, w, 1)
    assert np.array_equal(ids[: box["r"]["forest_count"]], k_ids)
    with pytest.raises(P.PhemError) as e:
        srv.wait(t)
    assert e.value.code == P.ENOTFOUND
    srv.close()


@pytest.mark.parametrize("rounds", [12, 2])
def test_recorded_batches_equal_oracle(P, rounds):
    """With record_rounds set, a capacity bucket's first batch runs eagerly
    and later batches of the bucket replay its recorded solve (host and
    device inputs mixed; the same 40 graphs submitted in four waves, each
    gathered into one batch by the batching window, so that buckets repeat);
    with too few rounds a replay falls back to the eager solve.  Either way
    every result equals the oracle's."""
    graphs = [host_graph(g, seed=7) for g in range(40)]
    srv = P.Server(mem_bytes=512 << 20, slots=2, max_graphs=64, record_rounds=rounds,
                   window_us=50000)
    want = [O.kruskal(n, u, v, w, 1) for (n, u, v, w) in graphs]
    for wave in range(4):
        out = [submit(srv, n, u, v, w, on_device=(g % 2 == wave % 2))
               for g, (n, u, v, w) in enumerate(graphs)]
        srv.drain()
        for g, (n, u, v, w) in enumerate(graphs):
            t, ids, lab = out[g]
            r = srv.wait(t)
            assert r["status"] == 0, r["error"]
            k_ids, k_lab = want[g]
            assert np.array_equal(ids[: r["forest_count"]], k_ids), (wave, g)
            assert np.array_equal(lab[:n], k_lab), (wave, g)
            assert r["weight_fx"] == O.weight_sum(k_ids, w, 1)[0], (wave, g)
    st = srv.stats()
    print(f"rounds={rounds}: {st['batches']} batches, {st['recorded_batches']} replayed, "
          f"{st['record_fallbacks']} fell back, {st['recordings']} recordings", flush=True)
    assert st["recorded_batches"] > 0 and st["recordings"] > 0
    if rounds == 2:
        assert st["record_fallbacks"] > 0
    srv.close()


def test_resident_streams_in_the_mix_equal_oracle(P):
    """Two resident streams fed interleaved batches (host and device inputs)
    while graph requests are served around them: every batch's result (the
    forest's edge count and exact weight after it) and the labels returned
    with it equal the oracle's Kruskal forest of every edge the stream has
    received so far; the graph requests equal the oracle too.  Then the
    stream memory budget, the batch-size limit and the close rules."""
    srv = P.Server(slots=3, max_batch_edges=1 << 20, wmode=P.W_F32)
    specs = [(1 << 12, 1 << 11, 11), (3000, 700, 12)]  # (n, batch edges, seed)
    sids = [srv.stream_open(n, mb) for n, mb, _ in specs]
    fed = [([], [], []) for _ in specs]
    expect = []   # (kind, ticket, payload)
    graphs = [host_graph(g, seed=8) for g in range(24)]
    gi = 0
    for b in range(8):
        for s, (n, mb, seed) in enumerate(specs):
            if b == 0:
                u, v, w = S.rmat(12, 4 * n, seed, positive=True) if s == 0 else \
                    S.uniform(n, mb, S.stream_seed(seed, 99), positive=True)
                u, v, w = u[:mb], v[:mb], w[:mb]
            else:
                u, v, w = S.uniform(n, mb if b % 3 else mb // 3, S.stream_seed(seed, b))
            for k, x in enumerate((u, v, w)):
                fed[s][k].append(x)
            lab = np.zeros(n, np.uint32)
            on_device = (b + s) % 2 == 1
            if on_device:
                t = srv.submit_batch(sids[s], dev(u), dev(v), dev(w), lab)
            else:
                t = srv.submit_batch(sids[s], u, v, w, lab)
            au, av, aw = (np.concatenate(x) for x in fed[s])
            expect.append(("stream", t, (s, n, au, av, aw, lab)))
        for _ in range(3):
            n, u, v, w = graphs[gi]
            t, ids, lab = submit(srv, n, u, v, w, on_device=gi % 2 == 0)
            expect.append(("graph", t, (n, u, v, w, ids, lab)))
            gi += 1
    for kind, t, pay in expect:
        r = srv.wait(t)
        assert r["status"] == 0, r["error"]
        if kind == "stream":
            s, n, au, av, aw, lab = pay
            k_ids, k_lab = O.kruskal(n, au, av, aw, 1)
            ws, exact = O.weight_sum(k_ids, aw, 1)
            assert r["forest_count"] == len(k_ids), (s, len(au))
            assert exact and r["weight_fx"] == ws and r["weight_exact"], (s, len(au))
            assert np.array_equal(lab, k_lab), (s, len(au))
        else:
            n, u, v, w, ids, lab = pay
            k_ids, k_lab = O.kruskal(n, u, v, w, 1)
            assert r["forest_count"] == len(k_ids)
            assert np.array_equal(ids[:len(k_ids)], k_ids) and np.array_equal(lab[:n], k_lab)
    st = srv.stats()
    assert st["stream_batches"] == 16 and st["streams_open"] == 2
    assert st["stream_edges"] == sum(sum(x.size for x in f[0]) for f in fed)
    assert st["completed"] == 16 + 24 and st["stream_lat_p99"] >= st["stream_lat_p50"] > 0
    # a batch beyond the stream's max_batch, forest ids asked of a stream
    u, v, w = S.uniform(3000, 701, 5)
    with pytest.raises(P.PhemError):
        srv.submit_batch(sids[1], u, v, w)
    with pytest.raises(P.PhemError) as e:
        srv.stream_close(12345)
    assert e.value.code == P.ENOTFOUND
    for sid in sids:
        srv.stream_close(sid)
    assert srv.stats()["streams_open"] == 0 and srv.stats()["stream_bytes"] == 0
    with pytest.raises(P.PhemError) as e:
        srv.submit_batch(sids[0], u, v, w)
    assert e.value.code == P.ENOTFOUND
    srv.close()
    # admission by footprint: a 4 MiB stream budget refuses a 2^20-vertex
    # stream (over 100 MB) and admits a 16-vertex one (1.8 MB, most of it
    # the solve's fixed counter area)
    srv = P.Server(slots=1, max_batch_edges=1 << 16, wmode=P.W_F32, stream_mem_mib=4)
    with pytest.raises(P.PhemError) as e:
        srv.stream_open(1 << 20, 1 << 10)
    assert e.value.code == P.EBUSY
    assert srv.stats()["refused"] == 1 and srv.stats()["streams_open"] == 0
    small = srv.stream_open(16, 8)
    assert srv.stats()["streams_open"] == 1
    srv.close()  # closes the stream with the server
    del small


def test_stream_batches_from_many_threads_apply_in_submission_order(P):
    """Four client threads submit batches to one resident stream at once.
    Tickets are issued under the server's lock, so ticket order is the
    submission order; every batch's result equals a stand-alone stream fed
    the batches in ticket order."""
    n, mb = 5000, 1500
    srv = P.Server(slots=3, max_batch_edges=1 << 18, wmode=P.W_F32)
    sid = srv.stream_open(n, mb)
    got = []
    lock = threading.Lock()

    def client(c):
        for b in range(10):
            u, v, w = S.uniform(n, mb - 100 * c, S.stream_seed(40 + c, b))
            t = srv.submit_batch(sid, u, v, w) if b % 2 else \
                srv.submit_batch(sid, dev(u), dev(v), dev(w))
            with lock:
                got.append((t, (u, v, w)))

    th = [threading.Thread(target=client, args=(c,)) for c in range(4)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    got.sort(key=lambda g: g[0])
    res = [srv.wait(t) for t, _ in got]
    ref = P.Stream(n, mb, wmode=P.W_F32)
    for (t, (u, v, w)), r in zip(got, res):
        ref.apply(dev(u), dev(v), dev(w))
        s = ref.stats()
        assert r["status"] == 0, r["error"]
        assert (r["forest_count"], r["weight_fx"]) == (s["mst_count"], s["weight_fx"]), t
    ref.close()
    srv.stream_close(sid)
    srv.close()
