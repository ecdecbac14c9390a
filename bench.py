"""Benchmark: Boruvka MST + union-find CC on RMAT n=2^24, m=2^28 on one MI355X.

One step = one minimum-spanning-forest solve (forest ids, exact weight,
min-id labels) plus one union-find connected-components solve (labels +
spanning forest) over the same device-resident graph.  Inputs are generated
on the device before the timed region (RMAT a,b,c,d = .57,.19,.19,.05, seed 2,
weights k * 2^-24 in (0,1]).

Prints one JSON line.  `value` = m / (time per step) in edges/s.
"""
import argparse
import gc
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "phem-graph_amd"))

PEAK_HBM_GBS = 8000.0

WORKLOADS = {
    # name: (scale, log2 m, seed)
    "rmat_n2^24_m2^28": (24, 28, 2),
    "rmat_n2^20_m2^24": (20, 24, 2),
}


def mst_kernel_bytes(st, m, n):
    """Algorithmic bytes per solve of each MST kernel, from the solve's log.
    Input passes read 12 B per input edge (u, v, w; the id is the index) and
    write 16 B (u', v', key) per edge they keep; a compaction round reads 16 B
    per alive edge and writes 16 B per survivor; flatten and labels read and
    write the 4 B parent / label of every vertex; the giant-edge emission
    reads the 8 B per-vertex minimum and writes 16 B per emitted edge."""
    ml, lr = st["m_log"], st["light_rounds"]
    R = len(ml)
    # survivors written by round r: the next round's alive count within a phase
    out = [ml[r + 1] if (r + 1 < R and r + 1 != lr) else 0 for r in range(R)]
    b = {"k_compact": sum(16 * ml[r] + 16 * out[r] for r in range(R)),
         "k_labels": 8 * n, "k_flatten": 8 * n, "k_giant_bits": 4 * n + n // 8}
    if st["tau"] == 0xFFFFFFFFFFFFFFFF:  # canonical: round 0 is sweep + compact<R0>
        b["k_sweep0"] = 12 * m
        b["k_compact<R0>"] = 12 * m + 16 * (ml[1] if R > 1 else 0)
        b["k_compact"] -= 16 * ml[0] + 16 * out[0]
    else:
        b["k_light"] = 12 * m + 16 * st["light_edges"]
        g = st.get("giant_edges", 0)
        b["k_heavy"] = 12 * m + 16 * (st["heavy_edges"] - g)
        b["k_giant_fold"] = 12 * n
        b["k_giant_emit"] = 8 * n + 16 * g
    return b


def b_actual(st, m, n):
    """Algorithmic bytes of the solve as run: every edge pass and every
    component-side pass (16 B per active component per round: best key and
    parent), plus 12 B per vertex for parent init, labels and the forest
    output."""
    kb = mst_kernel_bytes(st, m, n)
    return sum(kb.values()) - kb["k_labels"] + 16 * sum(st["c_log"]) + 12 * n


def b_mst(m, n, m_log, c_log):
    """Algorithmic bytes of a whole MST solve: 32 S_m - 20 m + 16 S_c + 12 n."""
    return 32 * sum(m_log) - 20 * m + 16 * sum(c_log) + 12 * n


def b_cc(m, n):
    return 8 * m + 12 * n


def host_cpus():
    """What the host offers this process: nproc (os.cpu_count), the CPUs in
    its affinity mask, the cgroup CPU quota (cpu.max, in CPUs; None when
    unlimited or unreadable) and the CPU model name."""
    info = {"nproc": os.cpu_count()}
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        info["affinity"] = info["nproc"]
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = round(int(q) / int(per), 2)
    except (OSError, ValueError):
        pass
    info["cgroup_cpus"] = quota
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    info["cpu_model"] = model
    return info


def cpu_baseline(u_np, v_np, w_np, n, threads, runs):
    """The CPU path on the same edges: MST by the OpenMP restatement of PBBS
    filter-Kruskal (MST/src/utils.cu:794-841; the reference's own file cannot
    be compiled here, it includes CUB), CC by the reference's own HAST
    host_spanning_tree (CC/host_spanning_tree.cu:67-92, built from its
    sources into oracle/_ref with its ParlayLib) when that build is present,
    else by the OpenMP restatement.  One warm-up, then the median of `runs`
    timed runs (the PBBS time_loop convention, common/time_loop.h:3-18).
    Returns (edges/s of MST + CC, MST s, CC s, CC kind)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    import numpy as np
    ref_cc = os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libref_hostcc.so"))
    if ref_cc:
        os.environ["PARLAY_NUM_THREADS"] = str(threads)  # read at the scheduler's start
        rec = np.ascontiguousarray((np.minimum(u_np, v_np).astype(np.uint64) << np.uint64(32))
                                   | np.maximum(u_np, v_np).astype(np.uint64))
    tm, tc = [], []
    for r in range(runs + 1):
        t0 = time.perf_counter()
        O.par_mst(n, u_np, v_np, w_np, 1, threads=threads)
        t1 = time.perf_counter()
        if ref_cc:
            cc_s, workers = O.ref_host_cc_time(n, rec)
        else:
            O.par_cc(n, u_np, v_np, threads=threads)
            cc_s = time.perf_counter() - t1
        print(f"# cpu run {r}: mst {t1 - t0:.2f} s cc {cc_s:.2f} s"
              + (f" (reference, {workers} ParlayLib workers)" if ref_cc else ""),
              file=sys.stderr, flush=True)
        if r:
            tm.append(t1 - t0)
            tc.append(cc_s)
    tm.sort(), tc.sort()
    mst_s, cc_s = tm[len(tm) // 2], tc[len(tc) // 2]
    return len(u_np) / (mst_s + cc_s), mst_s, cc_s, ("reference" if ref_cc else "port")


def serve_bench(P, torch, graphs_n, seed, slots, max_batch_edges, load, host_graphs=1024,
                record_rounds=12):
    """The serving mix (graphs_n independent graphs of the mix, device
    resident): one-at-a-time plain solves, then the serving loop saturated
    (all requests at once) and with Poisson arrivals at `load` x the saturated
    request rate, each with batches solved eagerly and replaying recorded
    solves (record_rounds rounds per phase), then the first host_graphs
    graphs from host memory (all at once, the staging rate; then Poisson
    arrivals, latencies including the host-to-device and result copies).
    Every request returns forest ids, labels and the exact weight to host
    memory; latency = done - arrival."""
    import numpy as np
    import phem_workloads as W

    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    graphs = [W.mix_graph(seed, g, device=dev) for g in range(graphs_n)]
    torch.cuda.synchronize()
    edges = sum(g[1].numel() for g in graphs)
    nmax = max(g[0] for g in graphs)
    mmax = max(g[1].numel() for g in graphs)
    print(f"# serving mix: {graphs_n} graphs, {edges} edges, generated in "
          f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    out_ids = [np.zeros(g[0], np.uint32) for g in graphs]
    out_lab = [np.zeros(g[0], np.uint32) for g in graphs]

    # plain: one graph at a time through phem_mst_solve, results to host
    solver = P.MstSolver(nmax, mmax, device=dev)
    h_ids = torch.empty(nmax, dtype=torch.int32, pin_memory=True)
    h_lab = torch.empty(nmax, dtype=torch.int32, pin_memory=True)
    lat_plain = []
    plain_w = []

    def plain(n, u, v, w):
        a = time.perf_counter()
        ids, lab, st = solver.solve(u, v, w, P.W_F32, n=n)
        k = st["mst_count"]
        h_ids[:k].copy_(ids[:k], non_blocking=True)
        h_lab[:n].copy_(lab[:n], non_blocking=True)
        torch.cuda.synchronize()
        lat_plain.append(time.perf_counter() - a)
        return st

    for g in graphs[:4]:
        plain(*g)
    lat_plain.clear()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for g in graphs:
        plain_w.append(plain(*g)["weight_fx"])
    t_plain = time.perf_counter() - t0
    print(f"# plain: {t_plain:.2f} s", file=sys.stderr, flush=True)
    del solver

    def run_served(srv, gl, outs, arrivals):
        srv.reset_stats()
        tick = []
        base = P.Server.clock()
        for i, (n, u, v, w) in enumerate(gl):
            if arrivals is not None:
                at = base + arrivals[i]
                while P.Server.clock() < at:
                    pass
            else:
                at = -1.0
            tick.append(srv.submit(n, u, v, w, outs[0][i], outs[1][i], at))
        res = [srv.wait(t) for t in tick]
        st = srv.stats()
        wall = st["last_done"] - st["first_arrival"]
        return res, st, wall

    def leg(st, wall, e, extra=None):
        d = {"edges_per_s": e / wall, "wall_s": round(wall, 4), "batches": st["batches"],
             "mean_batch_graphs": round(st["mean_batch_graphs"], 1),
             "lat_p50_ms": round(st["lat_p50"] * 1e3, 3),
             "lat_p99_ms": round(st["lat_p99"] * 1e3, 3),
             "recorded_batches": st["recorded_batches"],
             "record_fallbacks": st["record_fallbacks"], "recordings": st["recordings"]}
        d.update(extra or {})
        return d

    bad = 0
    rng = np.random.default_rng(seed)
    modes = {}
    for label, rr in (("eager", 0), ("recorded", record_rounds)):
        srv = P.Server(mem_bytes=0, slots=slots, max_graphs=4096,
                       max_batch_edges=max_batch_edges, wmode=P.W_F32, record_rounds=rr)
        run_served(srv, graphs, (out_ids, out_lab), None)  # warm-up: memory, recordings
        res, st, wall = run_served(srv, graphs, (out_ids, out_lab), None)
        bad += sum(r["status"] != 0 or r["weight_fx"] != plain_w[i] for i, r in enumerate(res))
        rate = load * graphs_n / wall  # requests per second
        arrivals = np.cumsum(rng.exponential(1.0 / rate, graphs_n))
        res_p, st_p, wall_p = run_served(srv, graphs, (out_ids, out_lab), arrivals)
        bad += sum(r["status"] != 0 or r["weight_fx"] != plain_w[i] for i, r in enumerate(res_p))
        srv.close()
        modes[label] = {"saturated": leg(st, wall, edges),
                        "poisson": leg(st_p, wall_p, edges,
                                       {"load": load, "rate_req_per_s": round(rate, 1)})}
        print(f"# served {label}: saturated {edges / wall / 1e9:.2f} G edges/s, poisson "
              f"p50 {st_p['lat_p50'] * 1e3:.2f} ms p99 {st_p['lat_p99'] * 1e3:.2f} ms",
              file=sys.stderr, flush=True)
    # the headline mode: the lower Poisson p99 among the modes within 3 % of
    # the best saturated rate (both modes are reported under "modes")
    top = max(modes[k]["saturated"]["edges_per_s"] for k in modes)
    near = [k for k in modes if modes[k]["saturated"]["edges_per_s"] >= 0.97 * top]
    best = min(near, key=lambda k: modes[k]["poisson"]["lat_p99_ms"])
    # resident streams in the mix: 4 streams (an RMAT 2^20 base of 2^22 edges,
    # then batches of 2^16 uniform edges) whose 64 batches arrive as a
    # Poisson stream over the span of the graph requests' Poisson arrivals
    # (the same rate as the Poisson leg); every stream's forest weight after
    # each batch is checked against a stand-alone stream fed the same batches
    streams_leg = serve_streams_leg(P, torch, graphs, (out_ids, out_lab), run_served,
                                    slots, max_batch_edges,
                                    record_rounds if best == "recorded" else 0,
                                    load * graphs_n / (edges / modes[best]["saturated"]["edges_per_s"]),
                                    rng)
    print(f"# served with streams: graphs p99 {streams_leg['graph_lat_p99_ms']:.2f} ms, "
          f"stream batches p50 {streams_leg['stream_lat_p50_ms']:.2f} ms "
          f"p99 {streams_leg['stream_lat_p99_ms']:.2f} ms", file=sys.stderr, flush=True)
    # host-resident inputs on a prefix of the mix: every latency includes the
    # staging into pinned memory, the host-to-device copies and the result
    # copies back; all at once (the staging rate), then Poisson arrivals
    hk = min(host_graphs, graphs_n)
    hg = [(n, u.cpu().numpy().view(np.uint32), v.cpu().numpy().view(np.uint32),
           w.cpu().numpy().view(np.uint32)) for n, u, v, w in graphs[:hk]]
    h_edges = sum(len(g[1]) for g in hg)
    srv = P.Server(mem_bytes=0, slots=slots, max_graphs=4096, max_batch_edges=max_batch_edges,
                   wmode=P.W_F32, record_rounds=record_rounds if best == "recorded" else 0)
    run_served(srv, hg, (out_ids, out_lab), None)  # warm-up
    res_h, st_h, wall_h = run_served(srv, hg, (out_ids, out_lab), None)
    bad += sum(r["status"] != 0 or r["weight_fx"] != plain_w[i] for i, r in enumerate(res_h))
    rate_h = load * hk / wall_h
    arr_h = np.cumsum(rng.exponential(1.0 / rate_h, hk))
    res_hp, st_hp, wall_hp = run_served(srv, hg, (out_ids, out_lab), arr_h)
    bad += sum(r["status"] != 0 or r["weight_fx"] != plain_w[i] for i, r in enumerate(res_hp))
    srv.close()
    lp = sorted(lat_plain)
    pct = lambda l, p: l[max(0, int(np.ceil(p * len(l))) - 1)]
    return {
        "workload": f"serve_mix_{graphs_n}", "graphs": graphs_n, "edges": edges,
        "inputs": "device resident; results copied to host memory",
        "plain": {"edges_per_s": edges / t_plain, "wall_s": round(t_plain, 4),
                  "lat_p50_ms": round(pct(lp, .5) * 1e3, 3),
                  "lat_p99_ms": round(pct(lp, .99) * 1e3, 3)},
        "served_saturated": modes[best]["saturated"],
        "served_poisson": modes[best]["poisson"],
        "served_mode": best,
        "modes": modes,
        "served_host_inputs": leg(st_h, wall_h, h_edges,
                                  {"graphs": hk, "edges": h_edges,
                                   "staged_GB_per_s": round(12 * h_edges / wall_h / 1e9, 2)}),
        "served_host_poisson": leg(st_hp, wall_hp, h_edges,
                                   {"graphs": hk, "load": load,
                                    "rate_req_per_s": round(rate_h, 1),
                                    "latency": "done - arrival, including staging, "
                                               "host-to-device and result copies"}),
        "served_with_streams": streams_leg,
        "speedup_saturated_vs_plain": round(modes[best]["saturated"]["edges_per_s"] * t_plain / edges, 2),
        "speedup_poisson_vs_plain": round(modes[best]["poisson"]["edges_per_s"] * t_plain / edges, 2),
        "slots": slots, "max_batch_edges": max_batch_edges, "record_rounds": record_rounds,
        "weights_mismatch_vs_plain": bad,
    }


def serve_streams_leg(P, torch, graphs, outs, run_served, slots, max_batch_edges,
                      record_rounds, rate, rng, n_streams=4, batches=16, batch_log2=16):
    """The graph mix with Poisson arrivals at `rate` requests/s, served beside
    n_streams resident streams (phem_server_stream_open; an RMAT 2^20 base of
    2^22 edges each, applied before the timed run) whose n_streams x batches
    batches of 2^batch_log2 uniform edges arrive as a Poisson stream over the
    same span, stream s taking every n_streams-th batch.  Reports the latency
    (done - arrival) of the graph requests and of the stream batches apart,
    and the aggregate edges/s of both; each stream's forest size and exact
    weight after every batch are compared with a stand-alone phem_stream fed
    the same batches."""
    import numpy as np
    import phem_workloads as W

    dev = torch.device("cuda:0")
    ns, nb, sb = n_streams, batches, 1 << batch_log2
    sn, sbase = 1 << 20, 1 << 22
    srv = P.Server(mem_bytes=0, slots=slots, max_graphs=4096,
                   max_batch_edges=max_batch_edges, wmode=P.W_F32,
                   record_rounds=record_rounds)
    sids = [srv.stream_open(sn, sbase) for _ in range(ns)]
    bases = [P.gen_rmat(20, sbase, 100 + s, positive=True, device=dev) for s in range(ns)]
    sbat = [[P.gen_uniform(sn, sb, W.stream_seed(200 + s, b), device=dev) for b in range(nb)]
            for s in range(ns)]
    torch.cuda.synchronize()
    for s in range(ns):
        r = srv.wait(srv.submit_batch(sids[s], *bases[s]))
        assert r["status"] == 0, r["error"]
    run_served(srv, graphs, outs, None)  # warm-up
    srv.reset_stats()
    g_arr = np.cumsum(rng.exponential(1.0 / rate, len(graphs)))
    s_arr = np.cumsum(rng.exponential(g_arr[-1] / (ns * nb), ns * nb))
    events = sorted([(t, 0, i) for i, t in enumerate(g_arr)] +
                    [(t, 1, j) for j, t in enumerate(s_arr)])
    tick_g, tick_s = [None] * len(graphs), [None] * (ns * nb)
    base = P.Server.clock()
    for t, kind, i in events:
        at = base + t
        while P.Server.clock() < at:
            pass
        if kind == 0:
            n, u, v, w = graphs[i]
            tick_g[i] = srv.submit(n, u, v, w, outs[0][i], outs[1][i], at)
        else:
            tick_s[i] = srv.submit_batch(sids[i % ns], *sbat[i % ns][i // ns], arrival=at)
    res_g = [srv.wait(t) for t in tick_g]
    res_s = [srv.wait(t) for t in tick_s]
    st = srv.stats()
    for sid in sids:
        srv.stream_close(sid)
    srv.close()
    bad = sum(r["status"] != 0 for r in res_g + res_s)
    # stand-alone streams fed the same batches in the same order
    same = bad == 0
    for s in range(ns):
        ref = P.Stream(sn, sbase, wmode=P.W_F32, device=dev)
        ref.apply(*bases[s])
        for b in range(nb):
            ref.apply(*sbat[s][b])
            rs = ref.stats()
            r = res_s[b * ns + s]
            same = same and r["forest_count"] == rs["mst_count"] and r["weight_fx"] == rs["weight_fx"]
        ref.close()
    lat = lambda rs: sorted(r["done"] - r["arrival"] for r in rs)
    pct = lambda l, p: l[max(0, int(np.ceil(p * len(l))) - 1)]
    lg, ls = lat(res_g), lat(res_s)
    e_g = sum(g[1].numel() for g in graphs)
    wall = st["last_done"] - st["first_arrival"]
    return {"streams": ns, "stream_vertices": sn, "stream_base_edges": sbase,
            "stream_batches": ns * nb, "stream_batch_edges": sb,
            "graphs": len(graphs), "rate_req_per_s": round(rate, 1),
            "edges_per_s": (e_g + ns * nb * sb) / wall,
            "graph_lat_p50_ms": round(pct(lg, .5) * 1e3, 3),
            "graph_lat_p99_ms": round(pct(lg, .99) * 1e3, 3),
            "stream_lat_p50_ms": round(pct(ls, .5) * 1e3, 3),
            "stream_lat_p99_ms": round(pct(ls, .99) * 1e3, 3),
            "failed": bad, "streams_equal_standalone": bool(same)}


def replay_bench(P, torch, scale, m_log2, seed, reps=30):
    """Recorded solve: one RMAT graph (n = 2^scale, 2^m_log2 edges) solved
    eagerly (phem_mst_solve) and replayed from its HIP graph
    (phem_mst_graph_launch + phem_mst_graph_stats); host time per solve,
    median of `reps`, and whether the replay's forest equals the eager one."""
    dev = torch.device("cuda:0")
    n, m = 1 << scale, 1 << m_log2
    u, v, w = P.gen_rmat(scale, m, seed, positive=True, device=dev)
    solver = P.MstSolver(n, m, device=dev)
    g = P.MstGraph(n, u, v, w, P.W_F32, rounds=0)

    def timed(f):
        for _ in range(3):
            f()
        ts = []
        for _ in range(reps):
            torch.cuda.synchronize()
            a = time.perf_counter()
            r = f()
            ts.append(time.perf_counter() - a)
        ts.sort()
        return ts[len(ts) // 2], r

    t_e, (ids, _, st) = timed(lambda: solver.solve(u, v, w, P.W_F32))
    t_g, (gids, _, gst) = timed(lambda: (g.launch(), g.result())[1])
    same = bool(torch.equal(ids, gids) and st["weight_fx"] == gst["weight_fx"])
    g.close()
    return {"workload": f"rmat_n2^{scale}_m2^{m_log2}", "eager_ms": round(t_e * 1e3, 4),
            "replay_ms": round(t_g * 1e3, 4), "rounds": st["rounds"],
            "replay_equals_eager": same}


def configs_bench(P, torch, reps=5):
    """The other single-graph configurations, timed like the headline step
    (device-resident inputs, HIP events, median of `reps` solves after a
    warm-up): the 4096 x 4096 tie-heavy grid (20 % of edges deleted, integer
    weights 1..16), the randomly permuted path on 2^22 vertices plus 2^20
    chords (deep rounds), both generated exactly as the parity tests
    generate them, and the headline RMAT graph with its vertex ids randomly
    permuted (the same edges under a layout without id locality).  Each
    reports MST and CC time, edges/s, and for the MST the algorithmic bytes
    of its own round log (32 S_m - 20 m + 16 S_c + 12 n, exact when the
    canonical schedule runs) over the time.  Last, the uniform graph on 2^16
    vertices with 2^20 edges, on the device and on the library's host backend
    (phem_cpu_msf, phem_cpu_cc, every CPU the process may use, median of 5
    after a warm-up)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import synth as S

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()

    def timed(fn):
        fn()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ts = []
        for _ in range(reps):
            ev[0].record(stream)
            out = fn()
            ev[1].record(stream)
            ev[1].synchronize()
            ts.append(ev[0].elapsed_time(ev[1]))
        return sorted(ts)[len(ts) // 2], out

    def one(name, n, du, dv, dw):
        m = du.numel()
        solver = P.MstSolver(n, m, device=dev)
        mst_ms, (_, _, st) = timed(lambda: solver.solve(du, dv, dw, P.W_F32))
        de = P.pack_edges(torch.minimum(du, dv), torch.maximum(du, dv))
        par = torch.empty(n, dtype=torch.int32, device=dev)
        lab = torch.empty(n, dtype=torch.int32, device=dev)
        cc_ms, _ = timed(lambda: P._check(P.lib().phem_cc_solve(
            n, P._p(de), m, P._p(par), None, P._p(lab), P._stream()), "phem_cc_solve"))
        can = solver.solve(du, dv, dw, P.W_F32, light_factor=0.0)[2]
        bm = b_mst(m, n, can["m_log"], can["c_log"])
        del solver, de, par, lab
        torch.cuda.empty_cache()
        print(f"# {name}: mst {mst_ms:.3f} ms cc {cc_ms:.3f} ms rounds {st['rounds']}",
              file=sys.stderr, flush=True)
        return {"n": n, "m": m, "mst_ms": round(mst_ms, 4), "mst_edges_per_s": m / (mst_ms * 1e-3),
                "cc_ms": round(cc_ms, 4), "cc_edges_per_s": m / (cc_ms * 1e-3),
                "rounds": st["rounds"], "schedule": "split" if st["tau"] != P.NO_EDGE else "canonical",
                "forest": st["mst_count"],
                "B_canonical_bytes": bm,
                "mst_roofline_frac_canonical_bytes": round(bm / (mst_ms * 1e-3) / 8e12, 4),
                "cc_roofline_frac": round(b_cc(m, n) / (cc_ms * 1e-3) / 8e12, 4)}

    out = {}
    u, v, w = S.grid(4096, 3)
    t = [torch.from_numpy(x.view(np.int32)).to(dev) for x in (u, v, w)]
    out["grid_4096_tie_heavy"] = one("grid 4096^2", 4096 * 4096, *t)
    del t
    u, v, w = S.path_chords(1 << 22, 1 << 20, 4)
    t = [torch.from_numpy(x.view(np.int32)).to(dev) for x in (u, v, w)]
    out["path_2^22+2^20_chords"] = one("path 2^22 + chords", 1 << 22, *t)
    del t
    # RMAT 2^24 / 2^28 (the headline graph) with vertex ids permuted
    du, dv, dw = P.gen_rmat(24, 1 << 28, 2, positive=True, device=dev)
    g = torch.Generator(device=dev)
    g.manual_seed(9)
    perm = torch.randperm(1 << 24, device=dev, generator=g, dtype=torch.int64).to(torch.int32)
    du = perm[du.long()]
    dv = perm[dv.long()]
    del perm
    out["rmat_n2^24_m2^28_permuted"] = one("rmat permuted", 1 << 24, du, dv, dw)
    del du, dv, dw
    torch.cuda.empty_cache()
    # C1: uniform n = 2^16, m = 2^20 on the GPU, and on the library's host
    # backend (phem_cpu_msf, phem_cpu_cc) with every CPU the process may use
    du, dv, dw = P.gen_uniform(1 << 16, 1 << 20, 1, device=dev)
    c1 = one("uniform 2^16 / 2^20", 1 << 16, du, dv, dw)
    hu, hv, hw = (x.cpu().numpy().view(np.uint32) for x in (du, dv, dw))
    threads = len(os.sched_getaffinity(0))

    def host_time(fn, reps=5):
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return sorted(ts)[len(ts) // 2]
    c1["host_backend"] = {
        "threads": threads,
        "msf_ms": round(host_time(lambda: P.cpu_msf(1 << 16, hu, hv, hw, P.W_F32, threads)) * 1e3, 3),
        "cc_ms": round(host_time(lambda: P.cpu_cc(1 << 16, hu, hv, threads)) * 1e3, 3)}
    out["c1_uniform_2^16_2^20"] = c1
    del du, dv, dw
    return out


def stream_bench(P, torch, scale, m0_log2, batches, batch_log2, seed):
    """Streaming forest: an RMAT graph (n = 2^scale, 2^m0_log2 edges) applied
    as the first batch, then `batches` batches of 2^batch_log2 uniform
    insertions (all generated on the device beforehand).  Per-batch latency is
    the host time of phem_stream_apply (which returns with the forest and the
    union-find labels updated).  At the end the resident forest is compared
    with a from-scratch solve over every edge inserted.  The same batches then
    go through the union-find alone (spanning forest and labels, no minimum
    forest), timed per batch the same way."""
    import numpy as np
    import phem_workloads as W

    dev = torch.device("cuda:0")
    n, m0, mb = 1 << scale, 1 << m0_log2, 1 << batch_log2
    base = P.gen_rmat(scale, m0, seed, device=dev)
    bat = [P.gen_uniform(n, mb, W.stream_seed(seed, b), device=dev) for b in range(batches)]
    torch.cuda.synchronize()
    st = P.Stream(n, m0, wmode=P.W_F32, device=dev)
    t0 = time.perf_counter()
    st.apply(*base)
    t_base = time.perf_counter() - t0
    lat, rounds = [], []
    # the garbage of earlier legs is collected now, not inside a batch
    gc.collect()
    gc.disable()
    try:
        t0 = time.perf_counter()
        for b in range(batches):
            a = time.perf_counter()
            st.apply(*bat[b])
            lat.append(time.perf_counter() - a)
            rounds.append(st.stats()["rounds"])
        t_all = time.perf_counter() - t0
    finally:
        gc.enable()
    slow = [{"batch": b, "ms": round(lat[b] * 1e3, 3), "rounds": rounds[b]}
            for b in sorted(range(batches), key=lambda b: -lat[b])[:3]]
    s = st.stats()
    fu, fv, fw, fg = st.forest()
    gid = fg.cpu().numpy()
    stream_labels = st.labels()
    st.close()

    # Union-find alone over the same batches: the reference's batched
    # spanning-forest loop (CC/ExternalSpanningTree.cu:133-218) with the parent
    # array and the union-find forest resident; a batch's latency runs until
    # its unions have finished on the device.
    rec_base = P.pack_edges(base[0], base[1])
    recs = [P.pack_edges(x[0], x[1]) for x in bat]
    cpar = torch.empty(n, dtype=torch.int32, device=dev)
    cfor = torch.empty(n, dtype=torch.int64, device=dev)
    P.cc_init(n, cpar, cfor)
    P.cc_union(n, rec_base, cpar, cfor)
    torch.cuda.synchronize()
    lat_cc = []
    gc.collect()
    gc.disable()
    try:
        t0 = time.perf_counter()
        for b in range(batches):
            a = time.perf_counter()
            P.cc_union(n, recs[b], cpar, cfor)
            torch.cuda.synchronize()
            lat_cc.append(time.perf_counter() - a)
        t_cc = time.perf_counter() - t0
    finally:
        gc.enable()
    clab = torch.empty(n, dtype=torch.int32, device=dev)
    P.cc_labels(n, cpar, clab)
    cc_forest = int((cfor != -1).sum())
    del recs, rec_base

    # every edge inserted, in insertion order -> one solve from scratch
    cat = [torch.cat([base[i]] + [x[i] for x in bat]) for i in range(3)]
    del bat
    ids, lab, rs = P.mst_solve(n, cat[0], cat[1], cat[2], P.W_F32)
    same = bool(np.array_equal(ids.cpu().numpy().view(np.uint32).astype(np.int64), gid)
                and rs["weight_fx"] == s["weight_fx"]
                and torch.equal(lab, stream_labels))
    # the union-find stream: the same labels, and a forest of n - c edges
    comps = int(torch.unique(lab).numel())
    cc_same = bool(torch.equal(clab, lab)) and cc_forest == n - comps
    lat.sort()
    lat_cc.sort()
    pct = lambda x, p: x[max(0, int(np.ceil(p * len(x))) - 1)]
    return {"workload": f"stream_rmat_n2^{scale}_m2^{m0_log2}+{batches}x2^{batch_log2}",
            "base_ms": round(t_base * 1e3, 3),
            "batch_edges_per_s": batches * mb / t_all,
            "batch_lat_p50_ms": round(pct(lat, .5) * 1e3, 3),
            "batch_lat_p99_ms": round(pct(lat, .99) * 1e3, 3),
            "slowest_batches": slow,
            "forest_edges": len(gid), "edges_inserted": s["edges_inserted"],
            "forest_equals_recompute": same,
            "union_find_only": {
                "what": "phem_cc_union per batch on the resident parent array "
                        "(spanning forest and labels; no minimum forest)",
                "batch_edges_per_s": batches * mb / t_cc,
                "batch_lat_p50_ms": round(pct(lat_cc, .5) * 1e3, 4),
                "batch_lat_p99_ms": round(pct(lat_cc, .99) * 1e3, 4),
                "labels_equal_recompute": cc_same}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="rmat_n2^24_m2^28", choices=list(WORKLOADS))
    ap.add_argument("--cpu-sample-log2", type=int, default=28,
                    help="edges of the CPU baseline sample (prefix of the stream)")
    ap.add_argument("--cpu-runs", type=int, default=5,
                    help="timed CPU baseline runs after one warm-up (median reported)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-mix run")
    ap.add_argument("--no-stream", action="store_true", help="skip the streaming run")
    ap.add_argument("--no-replay", action="store_true",
                    help="skip the recorded-solve (HIP graph) run")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the grid, path and permuted-RMAT runs")
    ap.add_argument("--serve-graphs", type=int, default=4096)
    ap.add_argument("--serve-slots", type=int, default=3)
    ap.add_argument("--serve-batch-log2", type=int, default=27,
                    help="edges per served batch (log2)")
    ap.add_argument("--serve-load", type=float, default=0.8,
                    help="Poisson arrival rate as a fraction of the saturated rate")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="run one device copy of this size before the timed region "
                         "(a known-bytes reference for hardware counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only (--gpus {args.gpus} requested); "
              "sharding a graph over several GPUs is not implemented", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import phem_native as P

    P.require_device()
    scale, lm, seed = WORKLOADS[args.workload]
    n, m = 1 << scale, 1 << lm
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    print(f"# generating {args.workload} on device", file=sys.stderr, flush=True)
    du, dv, dw = P.gen_rmat(scale, m, seed, positive=True, device=dev)
    de = P.pack_edges(du, dv)
    solver = P.MstSolver(n, m, device=dev)
    cc_par = torch.empty(n, dtype=torch.int32, device=dev)
    cc_lab = torch.empty(n, dtype=torch.int32, device=dev)
    cc_forest = torch.empty(n, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream()

    def step():
        ids, lab, st = solver.solve(du, dv, dw, P.W_F32)
        P._check(P.lib().phem_cc_solve(n, P._p(de), m, P._p(cc_par), P._p(cc_forest),
                                       P._p(cc_lab), P._stream()), "phem_cc_solve")
        return st

    if args.calib_copy_mib:
        src = torch.empty(args.calib_copy_mib << 20, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    for i in range(args.warmup):
        st = step()
        print(f"# warmup {i}: rounds {st['rounds']} forest {st['mst_count']}",
              file=sys.stderr, flush=True)
    torch.cuda.synchronize()

    # separate MST / CC phase times on their own (events on the library's stream)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    ev[0].record(stream)
    st = solver.solve(du, dv, dw, P.W_F32)[2]
    ev[1].record(stream)
    P._check(P.lib().phem_cc_solve(n, P._p(de), m, P._p(cc_par), P._p(cc_forest),
                                   P._p(cc_lab), P._stream()), "phem_cc_solve")
    ev[2].record(stream)
    torch.cuda.synchronize()
    mst_ms = ev[0].elapsed_time(ev[1])
    cc_ms = ev[1].elapsed_time(ev[2])

    # Per-kernel times: one untimed pass of `steps` steps with events around
    # every launch.  The timed region below records events only around the
    # dominant kernel's launches (its roofline figure), so that instrumenting
    # the other ~60 launches per step adds nothing to ms_per_step.
    P.prof_only(None)
    P.prof_reset()
    P.prof_enable(True)
    for _ in range(args.steps):
        st = step()
    torch.cuda.synchronize()
    P.prof_enable(False)
    prof = P.prof_read()
    dname = max(prof.items(), key=lambda kv: kv[1][1])[0]

    P.prof_only(dname)
    P.prof_reset()
    P.prof_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    P.prof_enable(False)
    P.prof_only(None)
    dprof = P.prof_read()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = m / (ms_per_step * 1e-3)

    # roofline of the dominant kernel (largest share of the profiled pass)
    kb = mst_kernel_bytes(st, m, n)
    sf = P.get_tuning()["cc_sample_factor"]
    m1 = min(m, int(sf * n)) if sf > 0 else m
    kb["k_cc_union"] = 8 * m1
    # two filtered launches per solve (records m1..3*m1, then the rest)
    kb["k_cc_union_filtered"] = 8 * (m - m1)
    kb["k_cc_labels"] = 8 * n
    dlaunch, dms = dprof[dname]  # the dominant kernel, inside the timed region
    per_launch_ms = dms / dlaunch
    launches_per_step = dlaunch / args.steps
    if dname in kb:
        bytes_per_launch = kb[dname] / launches_per_step
        achieved = bytes_per_launch / (per_launch_ms * 1e-3) / 1e9
    else:
        bytes_per_launch, achieved = None, None
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get(args.workload, {}).get(dname)
        except (OSError, ValueError):
            traffic = None
    roof = {"bound": "hbm", "kernel": dname,
            "achieved": None if achieved is None else round(achieved, 1),
            "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": None if achieved is None else round(achieved / PEAK_HBM_GBS, 4),
            "traffic": traffic,
            "traffic_from": "profiles/traffic.json: rocprofv3 FETCH_SIZE and WRITE_SIZE "
                            "runs of this workload, calibrated with a 1 GiB copy",
            "bytes_per_launch": bytes_per_launch,
            "avg_launch_ms": round(per_launch_ms, 4)}

    # canonical round log (untimed) for the formula 32 S_m - 20 m + 16 S_c + 12 n
    can = solver.solve(du, dv, dw, P.W_F32, light_factor=0.0)[2]
    torch.cuda.synchronize()
    bm, bc = b_mst(m, n, can["m_log"], can["c_log"]), b_cc(m, n)
    ba = b_actual(st, m, n)
    out = {
        "metric": "MST/CC edges/s at RMAT n=2^24 m=2^28; batched vs plain; p50/p99 latency",
        "value": value, "unit": "edges/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64", "data": "synthetic",
        "config": {"workload": args.workload, "n": n, "m": m,
                   "rmat_abcd": [0.57, 0.19, 0.19, 0.05], "weights": "k*2^-24 in (0,1]",
                   "step": "MST solve + union-find CC solve"},
        "mst": {"ms": round(mst_ms, 4), "edges_per_s": m / (mst_ms * 1e-3),
                "rounds": st["rounds"], "light_rounds": st["light_rounds"],
                "forest": st["mst_count"], "light_edges": st["light_edges"],
                "heavy_edges": st["heavy_edges"], "giant_edges": st["giant_edges"],
                "B_canonical_bytes": bm,
                "roofline_frac_canonical_bytes": round(bm / (mst_ms * 1e-3) / 8e12, 4),
                "B_actual_bytes": ba,
                "roofline_frac_actual_bytes": round(ba / (mst_ms * 1e-3) / 8e12, 4),
                "m_log": st["m_log"], "c_log": st["c_log"],
                "canonical_m_log": can["m_log"], "canonical_c_log": can["c_log"]},
        "cc": {"ms": round(cc_ms, 4), "edges_per_s": m / (cc_ms * 1e-3), "B_bytes": bc,
               "roofline_frac": round(bc / (cc_ms * 1e-3) / 8e12, 4)},
        "kernels_ms_per_step": {k: round(v[1] / args.steps, 4) for k, v in
                                sorted(prof.items(), key=lambda kv: -kv[1][1])},
        "roofline": roof,
    }

    # the CPU baseline runs after every GPU leg: the host thread pools it
    # starts (OpenMP, ParlayLib) would otherwise share the host cores with
    # the serving and streaming legs' host threads
    cpu_in = None
    if not args.no_cpu:
        k = min(m, 1 << args.cpu_sample_log2)
        cpu_in = [x[:k].cpu().numpy().view(np.uint32) for x in (du, dv, dw)]
        # the uniform n = 2^16, m = 2^20 graph, timed whole on the same CPU path
        cpu_c1 = [x.cpu().numpy().view(np.uint32)
                  for x in P.gen_uniform(1 << 16, 1 << 20, 1, device=dev)]
    if not args.no_serve:
        del du, dv, dw, de, solver, cc_par, cc_lab, cc_forest
        torch.cuda.empty_cache()
        out["serving"] = serve_bench(P, torch, args.serve_graphs, 5, args.serve_slots,
                                     1 << args.serve_batch_log2, args.serve_load)
        torch.cuda.empty_cache()
    if not args.no_stream:
        if args.no_serve:
            del du, dv, dw, de, solver, cc_par, cc_lab, cc_forest
            torch.cuda.empty_cache()
        out["streaming"] = stream_bench(P, torch, 22, 24, 256, 20, 4)
    if not args.no_replay:
        out["recorded_solve"] = replay_bench(P, torch, 16, 20, 1)
    if not args.no_configs:
        out["configs"] = configs_bench(P, torch)
    if cpu_in is not None:
        host = host_cpus()
        # every CPU the process may use, within its cgroup quota
        threads = host["affinity"] or 1
        if host["cgroup_cpus"]:
            threads = min(threads, max(1, int(host["cgroup_cpus"])))
        u_np, v_np, w_np = cpu_in
        k = len(u_np)
        print(f"# cpu baseline on {k} edges, {threads} threads, host {host}",
              file=sys.stderr, flush=True)
        eps, t_mst, t_cc, cc_kind = cpu_baseline(u_np, v_np, w_np, n, threads, args.cpu_runs)
        what = "all" if k == m else f"the first {k} of the"
        out["cpu_baseline"] = {
            "value": eps, "unit": "edges/s", "cores": threads,
            "kind": "port",
            "sample": f"{what} {m} edges of the same RMAT graph (n={n}), median of "
                      f"{args.cpu_runs} runs after a warm-up: MST by the OpenMP PBBS "
                      f"filter-Kruskal restatement {t_mst:.2f} s + CC by "
                      + ("the reference's HAST host_spanning_tree (ParlayLib)"
                         if cc_kind == "reference" else "the OpenMP HAST restatement")
                      + f" {t_cc:.2f} s",
            "parts": {"mst": {"kind": "port", "s": round(t_mst, 4),
                              "edges_per_s": k / t_mst},
                      "cc": {"kind": cc_kind, "s": round(t_cc, 4),
                             "edges_per_s": k / t_cc}},
            "host": host}
        c1_eps, c1_mst, c1_cc, c1_kind = cpu_baseline(*cpu_c1, 1 << 16, threads, args.cpu_runs)
        out["cpu_baseline"]["parts"]["c1_uniform_2^16_2^20"] = {
            "edges_per_s": c1_eps, "mst_s": round(c1_mst, 5), "cc_s": round(c1_cc, 5),
            "cc_kind": c1_kind}

    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
