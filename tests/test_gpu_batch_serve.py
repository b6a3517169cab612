"""Batched (many ligands per launch) and served (continuous batching) forms:
results must be identical to the per-ligand search, alone or in the mix."""
import time

import threading

import numpy as np
import pytest

import cdocker_amd as cdk
from cdocker_amd import workloads

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu]", *a, flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = cdk.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def prots(ctx):
    out = []
    for name in ("p48", "p64"):
        g = workloads.protein_grids(name)
        out.append((g, cdk.Protein(ctx, g["values"], g["radii"], g["spacing"], g["mid"], g["length"])))
    yield out
    for _, p in out:
        p.close()


def same(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


def test_batch_uneven_ligands_equal_per_ligand(ctx, prots):
    say("batch uneven ligands")
    g, prot = prots[1]
    w = workloads.config3(nlig=6, nquat=24, conf_range=(1, 4), atom_range=(15, 60))
    ligs = [cdk.Ligand(prot, L["charge"], L["epsilon"], L["radius"]) for L in w["ligands"]]
    b = cdk.Batch(ctx, prot)
    for lig, L in zip(ligs, w["ligands"]):
        b.add(lig, L["conformers"], L["quats"])
    b.run(cdk.ENGINE_FUSED)
    for j, (lig, L) in enumerate(zip(ligs, w["ligands"])):
        E, idx, mc = b.results(j)
        E1, idx1, mc1 = cdk.search(ctx, prot, lig, L["conformers"], L["quats"], engine=cdk.ENGINE_FUSED)
        assert same(E, E1) and np.array_equal(idx, idx1) and same(mc, mc1)
    b.close()


def test_chunking_does_not_change_results(ctx, prots):
    say("chunking invariance")
    g, prot = prots[0]
    w = workloads.config1()
    L = w["ligands"][0]
    lig = cdk.Ligand(prot, L["charge"], L["epsilon"], L["radius"])
    q = L["quats"][:, :100]
    E0, i0, _ = cdk.search(ctx, prot, lig, L["conformers"], q)
    ctx.set_option("fused_chunk", 7)
    ctx.set_option("mid_group", 3)
    E1, i1, _ = cdk.search(ctx, prot, lig, L["conformers"], q)
    ctx.set_option("fused_chunk", 0)
    ctx.set_option("mid_group", 0)
    assert same(E0, E1) and np.array_equal(i0, i1)


def test_overlapped_z_stage_equals_one_stream(ctx, prots):
    """The atom form's z-inverse stage on a second stream beside the next
    launch set (two T2 buffers, event-ordered) gives the same bits as the
    one-stream order, replayed as a hipGraph and eagerly, over many small
    launch sets so both buffers and every cross-stream wait are used."""
    say("overlapped z stage")
    g, prot = prots[1]
    w = workloads.config2(nconf=3, nquat=120)
    L = w["ligands"][0]
    lig = cdk.Ligand(prot, L["charge"], L["epsilon"], L["radius"])
    ctx.set_option("pipeline", 2)
    ctx.set_option("fused_chunk", 37)
    try:
        res = {}
        for ovl in (0, 1):
            for graph in (1, 0):
                ctx.set_option("overlap_z", ovl)
                ctx.set_option("use_graph", graph)
                for rep in range(2):  # a second replay of the same graph
                    res[(ovl, graph, rep)] = cdk.search(ctx, prot, lig, L["conformers"], L["quats"])
                assert ctx.get_option("last_pipeline") == 2
        E0, i0, m0 = res[(0, 1, 0)]
        for key, (E, i, m) in res.items():
            assert same(E, E0) and np.array_equal(i, i0) and np.array_equal(m, m0), key
    finally:
        ctx.set_option("overlap_z", 1)
        ctx.set_option("use_graph", 1)
        ctx.set_option("fused_chunk", 0)
        ctx.set_option("pipeline", -1)
        lig.close()


def _fixed_mid(ctx, grp, depth, poses):
    ctx.set_option("pipeline", 0)  # the schedules below are the plane form's
    ctx.set_option("mid_group", grp)
    ctx.set_option("mid_depth", depth)
    ctx.set_option("mid_poses", poses)


def _tuned_mid(ctx):
    _fixed_mid(ctx, 0, 0, 0)
    ctx.set_option("pipeline", -1)


def test_mid_schedule_does_not_change_results(ctx, prots):
    """Mid-stage group size, register ring depth and poses per protein load
    (what the autotuner picks per launch shape) only reorder work: results
    are bit-identical."""
    say("mid schedule invariance")
    g, prot = prots[1]
    w = workloads.config2(nconf=3, nquat=40)
    L = w["ligands"][0]
    lig = cdk.Ligand(prot, L["charge"], L["epsilon"], L["radius"])
    try:
        ctx.set_option("pipeline", 0)
        E0, i0, _ = cdk.search(ctx, prot, lig, L["conformers"], L["quats"])
        # group 5 and 3 split pose runs at group edges; 1 pose per group
        # leaves the run one pose long
        for grp, depth, poses in ((4, 2, 1), (8, 3, 1), (32, 2, 1), (5, 3, 1), (4, 2, 2), (8, 2, 4), (5, 2, 2),
                                  (3, 2, 4), (1, 2, 4), (32, 2, 4)):
            _fixed_mid(ctx, grp, depth, poses)
            E1, i1, _ = cdk.search(ctx, prot, lig, L["conformers"], L["quats"])
            assert ctx.get_option("last_mid_depth") == depth
            assert ctx.get_option("last_mid_poses") == poses
            assert same(E0, E1) and np.array_equal(i0, i1), (grp, depth, poses)
    finally:
        _tuned_mid(ctx)


def test_mid_pose_runs_in_mixed_batches(ctx, prots):
    """Runs of poses sharing a protein load stop at ligand boundaries and skip
    poses whose box does not fit: a batch of uneven ligands gives the same
    bits with 1, 2 or 4 poses per load."""
    say("mid pose runs, mixed batch")
    g, prot = prots[1]
    w = workloads.config3(nlig=5, nquat=7, conf_range=(1, 3), atom_range=(15, 60))
    ligs = [cdk.Ligand(prot, L["charge"], L["epsilon"], L["radius"]) for L in w["ligands"]]

    def run():
        b = cdk.Batch(ctx, prot)
        for lig, L in zip(ligs, w["ligands"]):
            b.add(lig, L["conformers"], L["quats"])
        b.run(cdk.ENGINE_FUSED)
        out = [b.results(j)[:2] for j in range(len(ligs))]
        b.close()
        return out

    try:
        _fixed_mid(ctx, 16, 2, 1)
        base = run()
        for grp, poses in ((16, 2), (16, 4), (6, 4), (2, 2)):
            _fixed_mid(ctx, grp, 2, poses)
            got = run()
            for (E0, i0), (E1, i1) in zip(base, got):
                assert same(E0, E1) and np.array_equal(i0, i1), (grp, poses)
    finally:
        _tuned_mid(ctx)


def test_server_alone_equals_in_mix(ctx, prots):
    say("server alone vs mix")
    rs = np.random.default_rng(11)
    reqs = []
    for k in range(6):
        na = int(rs.integers(15, 40))
        nc = int(rs.integers(1, 4))
        nq = int(rs.choice([16, 40]))
        L = workloads.ligand(500 + k, na, nc)
        L["quats"] = workloads.quaternion_sets(600 + k, nc, nq)
        L["protein_index"] = k % 2
        reqs.append(L)
    # alone: one request at a time through a server
    srv = cdk.Server(ctx, [p for _, p in prots], 16 << 30, 64)
    alone = []
    for L in reqs:
        rid = srv.submit(L["protein_index"], L["charge"], L["epsilon"], L["radius"], L["conformers"], L["quats"])
        alone.append(srv.result(rid))
    # mixed: all submitted at once, small launch sets so they interleave
    srv.close()
    srv = cdk.Server(ctx, [p for _, p in prots], 16 << 30, 24)
    ids = [srv.submit(L["protein_index"], L["charge"], L["epsilon"], L["radius"], L["conformers"], L["quats"])
           for L in reqs]
    mixed = [srv.result(i) for i in ids]
    st = srv.stats()
    say("server stats", st)
    assert st["requests_done"] == len(reqs)
    for a, m, L in zip(alone, mixed, reqs):
        assert same(a[0], m[0]) and np.array_equal(a[1], m[1]) and same(a[2], m[2])
        # and equal to a direct search
        lig = cdk.Ligand(prots[L["protein_index"]][1], L["charge"], L["epsilon"], L["radius"])
        E, idx, mc = cdk.search(ctx, prots[L["protein_index"]][1], lig, L["conformers"], L["quats"])
        assert same(E, m[0]) and np.array_equal(idx, m[1])
        assert m[3] > 0
    srv.close()


def test_workspace_within_budget_with_overlap(prots):
    """The launch-set working arrays, including the overlapped atom form's
    second T2 buffer, stay within the context's workspace budget, with the
    overlap on and off, and across searches that alternate the 64^3, 96^3 and
    128^3 grids in one context (one pool, re-carved per search)."""
    say("workspace budget")
    g, prot0 = prots[1]
    budget = 1 << 30
    c = cdk.Context(0, ws_budget_bytes=budget)
    prot = cdk.Protein(c, g["values"], g["radii"], g["spacing"], g["mid"], g["length"])
    w = workloads.config2(nconf=2, nquat=1500)
    L = w["ligands"][0]
    lig = cdk.Ligand(prot, L["charge"], L["epsilon"], L["radius"])
    big = []
    try:
        for ovl, pipeline in ((1, 2), (0, 2), (1, -1), (1, 0)):
            c.set_option("overlap_z", ovl)
            c.set_option("pipeline", pipeline)
            cdk.search(c, prot, lig, L["conformers"], L["quats"])
            used = c.get_option("workspace_bytes")
            say("overlap", ovl, "pipeline", pipeline, "workspace bytes", used, "budget", budget)
            assert 0 < used <= budget
        c.set_option("overlap_z", 1)
        c.set_option("pipeline", -1)
        for name in ("p96", "p128"):
            gb = workloads.protein_grids(name)
            pb = cdk.Protein(c, gb["values"], gb["radii"], gb["spacing"], gb["mid"], gb["length"])
            big.append((pb, cdk.Ligand(pb, L["charge"], L["epsilon"], L["radius"])))
        order = [(big[1], "p128"), ((prot, lig), "p64"), (big[0], "p96"), (big[1], "p128"), ((prot, lig), "p64")]
        for (pp, ll), name in order:
            E, _, _ = cdk.search(c, pp, ll, L["conformers"][:1], L["quats"][:1, :400])
            used = c.get_option("workspace_bytes")
            say(name, "workspace bytes", used, "budget", budget, "valid", int(np.isfinite(E).sum()))
            assert 0 < used <= budget and np.isfinite(E).sum() > 0
    finally:
        for pb, lb in big:
            lb.close()
            pb.close()
        lig.close()
        prot.close()
        c.close()


def test_server_releases_collected_requests(ctx, prots):
    """Many requests submitted and collected: the server holds none of them
    afterwards, and a collected id is unknown."""
    say("server releases requests")
    srv = cdk.Server(ctx, [prots[1][1]], 16 << 30, 256)
    L = workloads.ligand(920, 20, 2)
    q = workloads.quaternion_sets(921, 2, 30)
    for rnd in range(3):
        ids = [srv.submit(0, L["charge"], L["epsilon"], L["radius"], L["conformers"], q) for _ in range(20)]
        for i in ids:
            assert srv.result(i) is not None
        st = srv.stats()
        assert st["held_requests"] == 0 and st["requests_done"] == 20 * (rnd + 1), st
    with pytest.raises(cdk.CdkError):
        srv.result(ids[-1], wait_ms=0)
    # two callers waiting on the same request: one gets its results, the
    # other an error (the request was released under it), neither crashes
    for rnd in range(3):
        rid = srv.submit(0, L["charge"], L["epsilon"], L["radius"], L["conformers"], q)
        got = []

        def collect():
            try:
                got.append(("ok", srv.result(rid)))
            except cdk.CdkError as e:
                got.append(("err", str(e)))

        ths = [threading.Thread(target=collect) for _ in range(2)]
        for th in ths:
            th.start()
        for th in ths:
            th.join(60)
        kinds = sorted(k for k, _ in got)
        assert kinds == ["err", "ok"], got
    assert srv.stats()["held_requests"] == 0
    srv.close()


def test_server_budget_rejects_oversized(ctx, prots):
    say("server budget")
    ws = 8 << 30  # the context's default workspace budget
    with pytest.raises(cdk.CdkError) as ei:
        cdk.Server(ctx, [prots[1][1]], ws, 64)  # nothing left for requests
    assert ei.value.code == 6
    srv = cdk.Server(ctx, [prots[1][1]], ws + (64 << 10), 64)
    L = workloads.ligand(900, 30, 2)
    small = workloads.quaternion_sets(901, 2, 50)  # ~10 KB of state: admitted
    big = workloads.quaternion_sets(902, 2, 2000)  # ~420 KB: refused
    rid = srv.submit(0, L["charge"], L["epsilon"], L["radius"], L["conformers"], small)
    assert srv.result(rid) is not None
    with pytest.raises(cdk.CdkError) as ei:
        srv.submit(0, L["charge"], L["epsilon"], L["radius"], L["conformers"], big)
    assert ei.value.code == 6
    srv.close()


def test_server_poll_pending_then_done(ctx, prots):
    say("server poll")
    srv = cdk.Server(ctx, [prots[1][1]], 16 << 30, 256)
    L = workloads.ligand(910, 40, 4)
    q = workloads.quaternion_sets(911, 4, 500)
    rid = srv.submit(0, L["charge"], L["epsilon"], L["radius"], L["conformers"], q)
    r = srv.result(rid, wait_ms=0)
    t0 = time.time()
    while r is None and time.time() - t0 < 60:
        r = srv.result(rid, wait_ms=50)
    assert r is not None and np.isfinite(r[0]).any()
    srv.close()
