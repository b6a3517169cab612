"""End-to-end parity of the search (both engines) against the oracle."""
import numpy as np
import pytest

import cdocker_amd as cdk
from cdocker_amd import workloads
from oracle import oracle as orc
from tests.parity import check_search

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu]", *a, flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = cdk.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def p48(ctx):
    g = workloads.protein_grids("p48")
    prot = cdk.Protein(ctx, g["values"], g["radii"], g["spacing"], g["mid"], g["length"])
    yield g, prot
    prot.close()


@pytest.fixture(scope="module")
def p64(ctx):
    g = workloads.protein_grids("p64")
    prot = cdk.Protein(ctx, g["values"], g["radii"], g["spacing"], g["mid"], g["length"])
    yield g, prot
    prot.close()


def test_protein_spectrum_vs_numpy(p48):
    say("protein spectrum 48")
    g, prot = p48
    for c in (0, 8):
        S = prot.spectrum(c)
        ref = np.fft.rfftn(g["values"][c].astype(np.float64))
        assert np.abs(S - ref).max() / np.abs(ref).max() < 2e-6


@pytest.mark.parametrize("engine", [cdk.ENGINE_REFERENCE, cdk.ENGINE_FUSED])
def test_config1_subset_vs_oracle(ctx, p48, engine):
    """48^3 x 9, 24-atom ligand, 60 of the 500 quaternions."""
    say("cfg1 subset engine", engine)
    g, prot = p48
    w = workloads.config1()
    L = w["ligands"][0]
    lig = cdk.Ligand(prot, L["charge"], L["epsilon"], L["radius"])
    quats = L["quats"][:, :60]
    E, idx, mc = cdk.search(ctx, prot, lig, L["conformers"], quats, engine=engine)
    prob = orc.PoseProblem(g["values"], g["radii"], g["spacing"], L["charge"], L["epsilon"], L["radius"])
    stats = check_search(prob, L["conformers"], quats, E, idx, mc)
    say("cfg1 engine", engine, stats)
    assert stats.get("exact", 0) >= 55


def test_engines_agree_and_are_deterministic(ctx, p48):
    say("engines agree")
    g, prot = p48
    w = workloads.config1()
    L = w["ligands"][0]
    lig = cdk.Ligand(prot, L["charge"], L["epsilon"], L["radius"])
    quats = L["quats"][:, :200]
    Er, ir, mr = cdk.search(ctx, prot, lig, L["conformers"], quats, engine=cdk.ENGINE_REFERENCE)
    Ef, jf, mf = cdk.search(ctx, prot, lig, L["conformers"], quats, engine=cdk.ENGINE_FUSED)
    Ef2, jf2, mf2 = cdk.search(ctx, prot, lig, L["conformers"], quats, engine=cdk.ENGINE_FUSED)
    assert np.array_equal(Ef.view(np.uint32), Ef2.view(np.uint32)) and np.array_equal(jf, jf2)
    assert np.array_equal(mr, mf)
    fin = np.isfinite(Er)
    assert np.array_equal(fin, np.isfinite(Ef))
    assert np.allclose(Er[fin], Ef[fin], rtol=1e-3, atol=2e-3)
    assert (ir == jf).all(axis=-1).mean() > 0.95


def test_config2_subset_vs_oracle(ctx, p64):
    """64^3 x 9, 40-atom ligand: 2 conformers x 12 quaternions per engine."""
    g, prot = p64
    w = workloads.config2(nconf=2, nquat=12)
    L = w["ligands"][0]
    lig = cdk.Ligand(prot, L["charge"], L["epsilon"], L["radius"])
    prob = orc.PoseProblem(g["values"], g["radii"], g["spacing"], L["charge"], L["epsilon"], L["radius"])
    for engine in (cdk.ENGINE_REFERENCE, cdk.ENGINE_FUSED):
        say("cfg2 subset engine", engine)
        E, idx, mc = cdk.search(ctx, prot, lig, L["conformers"], L["quats"], engine=engine)
        stats = check_search(prob, L["conformers"], L["quats"], E, idx, mc)
        say("cfg2 engine", engine, stats)
        assert stats.get("exact", 0) >= 20


RAMP = np.array([[0.25, -0.5, 0.125, 1.0], [-0.125, 0.25, 0.5, -2.0], [0.5, 0.125, -0.25, 0.5],
                 [0.0625, -0.25, 0.375, 0.0], [-0.375, 0.5, 0.0625, 3.0], [0.125, 0.125, -0.5, -1.0],
                 [0.25, 0.375, 0.25, 0.25], [-0.5, -0.0625, 0.125, 2.0], [0.375, -0.125, -0.375, -0.5]])


@pytest.mark.parametrize("N", [64, 96])
def test_linear_ramp_protein_closed_form(ctx, N):
    """Protein channels that are linear ramps P_c(i, j, k) = a_c i + b_c j +
    g_c k + d_c: trilinear spreading reproduces them exactly, so a pose's
    energy at translation t is sum_c sum_{a in c} w_a P_c(u_a + t), u_a the
    atom's continuous grid position, linear in t.  Every engine form (the
    reference dataflow, the plane, column and atom forms, all three product
    forms) must give that energy at its translation, and the argmin must be
    the window corner the slopes' signs pick.  No oracle involved."""
    h = 0.5
    i, j, k = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    vals = np.stack([(a * i + b * j + g * k + d).astype(np.float32) for a, b, g, d in RAMP])
    radii = np.array([1.2, 1.5, 1.7, 1.8, 1.9, 2.0, 2.1, 2.2], np.float32)
    mid = np.zeros(3, np.float32)
    length = np.full(3, N * h, np.float32)
    prot = cdk.Protein(ctx, vals, radii, h, mid, length)
    L = workloads.ligand(971, 30, 2)
    lig = cdk.Ligand(prot, L["charge"], L["epsilon"], L["radius"])
    nu, used, ari = lig.channels()
    quats = workloads.quaternion_sets(972, 2, 6)
    forms = [("reference", cdk.ENGINE_REFERENCE, -1, 2), ("planes", cdk.ENGINE_FUSED, 0, 2),
             ("atoms x3", cdk.ENGINE_FUSED, 2, 2), ("atoms x2", cdk.ENGINE_FUSED, 2, 0),
             ("atoms fp32", cdk.ENGINE_FUSED, 2, 1)]
    if N == 64:
        forms.append(("columns", cdk.ENGINE_FUSED, 1, 2))
    # channel of each atom: used vdW channels in used order, elec last
    chan = np.full(len(ari), -1)
    for c, r in enumerate(used):
        chan[ari == r] = c
    sq = np.sqrt(L["epsilon"].astype(np.float64))
    q64 = L["charge"].astype(np.float64)
    try:
        for label, engine, pipe, products in forms:
            ctx.set_option("pipeline", pipe)
            ctx.set_option("atom_products", products)
            E, idx, mc = cdk.search(ctx, prot, lig, L["conformers"], quats, engine=engine)
            if engine == cdk.ENGINE_FUSED:
                assert ctx.get_option("last_pipeline") == pipe
            worst = 0.0
            for c in range(E.shape[0]):
                for qi in range(E.shape[1]):
                    rot32 = orc.rotate(quats[c, qi], L["conformers"][c])[0]
                    rot = rot32.astype(np.float64)
                    u = (rot - mc[c, qi].astype(np.float64)) / h
                    # window in FP32 as NoMini:446-450 (h = 0.5: the division is exact)
                    span = (rot32.max(axis=0) - rot32.min(axis=0)).astype(np.float32)
                    win = N - (span / np.float32(h)).astype(np.int64) - 2
                    slope = np.zeros(3)
                    E_t = 0.0
                    for ch in range(nu + 1):
                        sel = chan == ch if ch < nu else np.ones(len(ari), bool)
                        w = sq[sel] if ch < nu else q64
                        coef = RAMP[used[ch]] if ch < nu else RAMP[8]
                        W = w.sum()
                        Wu = (w[:, None] * u[sel]).sum(axis=0)
                        slope += coef[:3] * W
                        E_t += float(coef[:3] @ (Wu + idx[c, qi] * W) + coef[3] * W)
                    scale = np.abs(slope).sum() * N + abs(E_t)
                    err = abs(float(E[c, qi]) - E_t) / scale
                    worst = max(worst, err)
                    assert err <= 1e-5, (label, c, qi, float(E[c, qi]), E_t)
                    corner = [0 if slope[d] > 0 else win[d] - 1 for d in range(3)]
                    for d in range(3):
                        # a slope well above the FP32 transforms' noise picks the corner
                        if abs(slope[d]) > 1e-4 * scale:
                            assert idx[c, qi, d] == corner[d], (label, c, qi, d, list(idx[c, qi]), corner)
            say("ramp protein", N, label, "max scaled error %.2e" % worst)
    finally:
        ctx.set_option("pipeline", -1)
        ctx.set_option("atom_products", 2)
        lig.close()
        prot.close()


def test_edge_sizes(ctx, p64):
    """Edge sizes: a ligand of the maximum 512 atoms (compact, so its box fits
    the 64^3 grid) through the atom form and the reference dataflow against
    the FP64 oracle over the whole window; 513 atoms refused; searches with no
    rotations or no conformers return empty results; a batch whose middle job
    has no rotations keeps the other jobs' results."""
    from tests.parity import check_sampled
    g, prot = p64
    rs = np.random.default_rng(981)
    na = 512
    L = workloads.ligand(981, na, 1)
    d = rs.normal(size=(na, 3))
    x = (7.0 * d / np.linalg.norm(d, axis=1, keepdims=True) * rs.uniform(0, 1, (na, 1)) ** (1 / 3)).astype(np.float32)
    lig = cdk.Ligand(prot, L["charge"], L["epsilon"], L["radius"])
    prob = orc.PoseProblem(g["values"], g["radii"], g["spacing"], L["charge"], L["epsilon"], L["radius"])
    quats = workloads.quaternion_sets(982, 1, 3)
    try:
        for engine, pipe in ((cdk.ENGINE_FUSED, 2), (cdk.ENGINE_REFERENCE, -1)):
            ctx.set_option("pipeline", pipe)
            E, idx, mc = cdk.search(ctx, prot, lig, x[None], quats, engine=engine)
            st = check_sampled(prob, x[None], np.zeros(3, np.int32), quats[0], E[0], idx[0], mc[0],
                               "512 atoms engine %d" % engine)
            say("512-atom ligand, engine", engine, st)
            assert st["valid"] == 3 and st["exact"] + st["near_ties"] == 3
        # no rotations, no conformers
        E, idx, mc = cdk.search(ctx, prot, lig, x[None], np.zeros((1, 0, 4), np.float32))
        assert E.shape == (1, 0) and idx.shape == (1, 0, 3)
        E, idx, mc = cdk.search(ctx, prot, lig, np.zeros((0, na, 3), np.float32), np.zeros((0, 5, 4), np.float32))
        assert E.shape == (0, 5)
    finally:
        ctx.set_option("pipeline", -1)
        lig.close()
    with pytest.raises(cdk.CdkError):
        L2 = workloads.ligand(983, 513, 1)
        cdk.Ligand(prot, L2["charge"], L2["epsilon"], L2["radius"])
    # a batch with an empty middle job
    Ls = [workloads.ligand(984 + i, 20, 2) for i in range(3)]
    ligs = [cdk.Ligand(prot, Li["charge"], Li["epsilon"], Li["radius"]) for Li in Ls]
    qs = [workloads.quaternion_sets(990 + i, 2, n) for i, n in enumerate((5, 0, 7))]
    try:
        b = cdk.Batch(ctx, prot)
        for lg, Li, q in zip(ligs, Ls, qs):
            b.add(lg, Li["conformers"], q)
        b.run(cdk.ENGINE_FUSED)
        for j in (0, 2):
            E, idx, mc = b.results(j)
            E1, i1, m1 = cdk.search(ctx, prot, ligs[j], Ls[j]["conformers"], qs[j])
            assert np.array_equal(E.view(np.uint32), E1.view(np.uint32)) and np.array_equal(idx, i1)
        assert b.results(1)[0].shape == (2, 0)
        b.close()
    finally:
        for lg in ligs:
            lg.close()


def test_shared_quaternions_and_empty_windows(ctx, p48):
    say("shared quats / empty windows")
    g, prot = p48
    x, q, e, r = cdk.synth_ligand(55, 20)
    big = x * np.float32(4.0)  # too large for most rotations -> empty windows
    lig = cdk.Ligand(prot, q, e, r)
    rng = cdk.Rng(4)
    quats = cdk.QuaternionUniformSampling(rng, 16)
    confs = np.stack([x, big])
    for engine in (cdk.ENGINE_REFERENCE, cdk.ENGINE_FUSED):
        E, idx, mc = cdk.search(ctx, prot, lig, confs, quats, engine=engine)
        prob = orc.PoseProblem(g["values"], g["radii"], g["spacing"], q, e, r)
        stats = check_search(prob, confs, quats, E, idx, mc)
        assert stats.get("empty", 0) >= 1


def test_rank_poses_rebuild(ctx, p48):
    say("rank poses")
    g, prot = p48
    w = workloads.config1()
    L = w["ligands"][0]
    lig = cdk.Ligand(prot, L["charge"], L["epsilon"], L["radius"])
    quats = L["quats"][0, :40]
    E, idx, mc = cdk.search(ctx, prot, lig, L["conformers"], quats, engine=cdk.ENGINE_FUSED)
    rot = ctx.Rotate(quats, L["conformers"][0])
    rank, pose = cdk.rank_poses(prot, rot, E[0], idx[0], mc[0], 5)
    assert rank.tolist() == cdk.sort_index(E[0])[:5].tolist()
    gmin = prot.grid_min()
    for i, rq in enumerate(rank):
        shift = gmin - mc[0, rq].astype(np.float64) + (idx[0, rq].astype(np.float32) * np.float32(prot.spacing))
        assert np.allclose(pose[i], rot[rq].astype(np.float64) + shift, atol=1e-9)


@pytest.mark.parametrize("variant", [0, 1])
def test_plane_stage_forms_vs_oracle(ctx, p48, p64, variant):
    """Both forms of the forward plane stage (block-wide LDS Stockham FFTs,
    register four-step FFTs) and of the inverse plane stage (block-wide LDS
    FFTs per plane, register FFTs over 4 planes) against the oracle."""
    ctx.set_option("pipeline", 0)
    ctx.set_option("fwd_variant", variant)
    ctx.set_option("inv_variant", variant)
    try:
        for (g, prot), w, nq in ((p48, workloads.config1(), 40), (p64, workloads.config2(nconf=1, nquat=10), 10)):
            say("forward form", variant, "grid", g["n"])
            L = w["ligands"][0]
            lig = cdk.Ligand(prot, L["charge"], L["epsilon"], L["radius"])
            quats = L["quats"][:1, :nq]
            E, idx, mc = cdk.search(ctx, prot, lig, L["conformers"][:1], quats)
            assert ctx.get_option("last_fwd_variant") == variant
            assert ctx.get_option("last_inv_variant") == variant
            prob = orc.PoseProblem(g["values"], g["radii"], g["spacing"], L["charge"], L["epsilon"], L["radius"])
            stats = check_search(prob, L["conformers"][:1], quats, E, idx, mc)
            say("forward form", variant, stats)
            assert stats.get("exact", 0) >= nq - 2
    finally:
        ctx.set_option("pipeline", -1)
        ctx.set_option("fwd_variant", -1)
        ctx.set_option("inv_variant", -1)


@pytest.mark.parametrize("pipeline", [1, 2])
def test_column_and_atom_forms_vs_oracle(ctx, p64, pipeline):
    """Column form at 64^3 (forward z transform only; y and x transforms per
    kz plane in the mid stage; z C2R last) and atom form (ligand spectra per
    kz plane from the atoms on the matrix cores; then as the column form)
    against the oracle, for ligands whose boxes fit in N/2 and ligands whose
    boxes do not; and the same bits for every mid-stage group size."""
    g, prot = p64
    ctx.set_option("pipeline", pipeline)
    try:
        worst = 0
        for seed, na, nq in ((901, 15, 8), (902, 40, 8), (903, 90, 6)):
            L = workloads.ligand(seed, na, 1)
            quats = workloads.quaternion_sets(seed + 1, 1, nq)
            lig = cdk.Ligand(prot, L["charge"], L["epsilon"], L["radius"])
            b = cdk.Batch(ctx, prot)
            b.add(lig, L["conformers"], quats)
            b.run(cdk.ENGINE_FUSED)
            box, _ = b.geometry(0)
            b.close()
            worst = max(worst, int(box.max()))
            say("pipeline", pipeline, "ligand", na, "atoms, largest box", int(box.max()))
            E, idx, mc = cdk.search(ctx, prot, lig, L["conformers"], quats)
            assert ctx.get_option("last_pipeline") == pipeline
            prob = orc.PoseProblem(g["values"], g["radii"], g["spacing"], L["charge"], L["epsilon"], L["radius"])
            stats = check_search(prob, L["conformers"], quats, E, idx, mc)
            say("pipeline", pipeline, na, stats)
            assert stats.get("exact", 0) >= nq - 1
            for grp in (1, 3, 8):
                ctx.set_option("mid_group", grp)
                E1, i1, _ = cdk.search(ctx, prot, lig, L["conformers"], quats)
                assert ctx.get_option("last_mid_group") == grp
                assert np.array_equal(E1.view(np.uint32), E.view(np.uint32)) and np.array_equal(i1, idx), grp
            ctx.set_option("mid_group", 0)
            lig.close()
        assert worst > 32  # the unpruned transforms ran too
    finally:
        ctx.set_option("mid_group", 0)
        ctx.set_option("pipeline", -1)


def test_autotuned_choice_is_sticky_and_reproducible(ctx, p64):
    say("autotuner")
    g, prot = p64
    w = workloads.config2(nconf=3, nquat=50)
    L = w["ligands"][0]
    lig = cdk.Ligand(prot, L["charge"], L["epsilon"], L["radius"])
    E0, i0, _ = cdk.search(ctx, prot, lig, L["conformers"], L["quats"])
    v = ctx.get_option("last_fwd_variant")
    vi = ctx.get_option("last_inv_variant")
    assert v in (0, 1) and vi in (0, 1) and ctx.get_option("last_mid_group") in (1, 2, 4, 8, 16, 32)
    pipe = ctx.get_option("last_pipeline")
    assert pipe in (0, 1, 2)
    # a different batch shape keeps the forward form of this grid size
    E1, i1, _ = cdk.search(ctx, prot, lig, L["conformers"][:1], L["quats"][:1, :7])
    assert ctx.get_option("last_fwd_variant") == v and ctx.get_option("last_inv_variant") == vi
    assert ctx.get_option("last_pipeline") == pipe
    assert np.array_equal(E1.view(np.uint32), E0[:1, :7].view(np.uint32)) and np.array_equal(i1, i0[:1, :7])
    ctx.set_option("fwd_variant", v)
    ctx.set_option("inv_variant", vi)
    E2, i2, _ = cdk.search(ctx, prot, lig, L["conformers"], L["quats"])
    ctx.set_option("fwd_variant", -1)
    ctx.set_option("inv_variant", -1)
    assert np.array_equal(E2.view(np.uint32), E0.view(np.uint32)) and np.array_equal(i2, i0)


def test_atom_form_large_ligand(ctx, p64):
    """A 300-atom ligand (600 channel atoms: more than the atom form stages at
    once, so it stages one channel at a time) against the oracle and the
    plane form, within the parity tolerance."""
    g, prot = p64
    rs = np.random.default_rng(931)
    L = workloads.ligand(931, 300, 1)
    # compact coordinates: a ball of radius 6 A keeps the box well inside 64^3
    d = rs.normal(size=(300, 3))
    x = 6.0 * d / np.linalg.norm(d, axis=1, keepdims=True) * rs.uniform(0, 1, (300, 1)) ** (1 / 3)
    confs = x[None].astype(np.float32)
    quats = workloads.quaternion_sets(932, 1, 3)
    lig = cdk.Ligand(prot, L["charge"], L["epsilon"], L["radius"])
    prob = orc.PoseProblem(g["values"], g["radii"], g["spacing"], L["charge"], L["epsilon"], L["radius"])
    try:
        res = {}
        for pipeline in (2, 0):
            ctx.set_option("pipeline", pipeline)
            E, idx, mc = cdk.search(ctx, prot, lig, confs, quats)
            assert ctx.get_option("last_pipeline") == pipeline
            stats = check_search(prob, confs, quats, E, idx, mc)
            say("300-atom ligand, pipeline", pipeline, stats)
            assert stats.get("exact", 0) >= 2
            res[pipeline] = (E, idx)
        # both forms within the energy tolerance of each other
        Ea, Ep = res[2][0], res[0][0]
        assert np.all(np.abs(Ea - Ep) <= 1e-4 * np.maximum(np.abs(Ep), 1.0))
    finally:
        ctx.set_option("pipeline", -1)
        lig.close()


@pytest.mark.parametrize("name", ["p64", "p96", "p128"])
def test_atom_products_all_forms(ctx, name):
    """The atom form's three product forms (FP16 matrix cores on exact
    three-part operand splits, the default; two-part splits; FP32 matrix
    cores) against the oracle, each within the parity tolerance with a scaled
    error well under it, and the option in effect (the forms round
    differently).  Ligands with tiny weights (charges down to 1e-6, whose
    low FP16 parts would be subnormal without the power-of-two scaling) too."""
    g = workloads.protein_grids(name)
    prot = cdk.Protein(ctx, g["values"], g["radii"], g["spacing"], g["mid"], g["length"])
    ctx.set_option("pipeline", 2)
    try:
        assert ctx.get_option("atom_products") == 2
        with pytest.raises(cdk.CdkError):
            ctx.set_option("atom_products", 3)
        assert ctx.get_option("atom_products") == 2
        for seed, na, nq, tiny in ((941, 23, 5, False), (942, 60, 4, False), (943, 30, 4, True)):
            L = workloads.ligand(seed, na, 1)
            charge, eps = L["charge"], L["epsilon"]
            if tiny:
                charge = (charge * np.float32(1e-4) * np.logspace(-2, 0, na).astype(np.float32)).astype(np.float32)
                eps = (eps * np.float32(1e-6)).astype(np.float32)
            quats = workloads.quaternion_sets(seed + 1, 1, nq)
            lig = cdk.Ligand(prot, charge, eps, L["radius"])
            prob = orc.PoseProblem(g["values"], g["radii"], g["spacing"], charge, eps, L["radius"])
            res = {}
            for products in (2, 0, 1):
                ctx.set_option("atom_products", products)
                E, idx, mc = cdk.search(ctx, prot, lig, L["conformers"], quats)
                assert ctx.get_option("last_pipeline") == 2
                stats = check_search(prob, L["conformers"], quats, E, idx, mc,
                                     label="%s products %d ligand %d" % (name, products, seed))
                say(name, "atom products", products, "ligand", na, "tiny weights" if tiny else "", stats)
                assert stats.get("exact", 0) >= nq - 1
                assert stats["max_scaled_err"] <= 1e-5
                res[products] = E
            assert not np.array_equal(res[2].view(np.uint32), res[1].view(np.uint32))
            lig.close()
    finally:
        ctx.set_option("atom_products", 2)
        ctx.set_option("pipeline", -1)
        prot.close()


@pytest.mark.parametrize("name", ["p64", "p96", "p128"])
def test_atom_form_channel_padding_edges(ctx, name):
    """Channel lists padded to multiples of 8 entries: a one-atom ligand (7
    pad slots per channel), one vdW channel of exactly 8 atoms (no pad) and
    one of 9 (a second step with one atom and 7 pads), against the oracle."""
    g = workloads.protein_grids(name)
    prot = cdk.Protein(ctx, g["values"], g["radii"], g["spacing"], g["mid"], g["length"])
    ctx.set_option("pipeline", 2)
    try:
        for seed, na, rad in ((951, 1, None), (952, 8, 1.7), (953, 9, 1.7), (954, 17, None)):
            L = workloads.ligand(seed, na, 1)
            radius = np.full(na, rad, np.float32) if rad is not None else L["radius"]
            quats = workloads.quaternion_sets(seed + 1, 1, 4)
            lig = cdk.Ligand(prot, L["charge"], L["epsilon"], radius)
            nu, used, ari = lig.channels()
            counts = [int((ari == u).sum()) for u in used] + [na]
            say("padding edges", name, "atoms", na, "channel counts", counts)
            if rad is not None:
                assert counts == [na, na]
            E, idx, mc = cdk.search(ctx, prot, lig, L["conformers"], quats)
            assert ctx.get_option("last_pipeline") == 2
            prob = orc.PoseProblem(g["values"], g["radii"], g["spacing"], L["charge"], L["epsilon"], radius)
            stats = check_search(prob, L["conformers"], quats, E, idx, mc)
            say("padding edges", name, na, stats)
            assert stats.get("exact", 0) + stats.get("empty", 0) >= 3
            lig.close()
    finally:
        ctx.set_option("pipeline", -1)
        prot.close()


@pytest.mark.parametrize("name", ["p96", "p128"])
def test_atom_form_96_128_vs_oracle(ctx, name):
    """Atom form at 96^3 and 128^3 (nine or sixteen 32 x 32 matrix-core blocks
    per kz plane, at 96 the twiddle index reduced mod 96; the inverse y
    transforms in the forward kernel, the window's y rows through HBM to the
    x-z kernel's 96- or 128-point transforms and argmin) against the oracle,
    and the same bits for every mid-stage group size.  The
    300-atom ligand's padded channel lists exceed the 512 staged entries, so
    the forward kernel stages them one channel at a time."""
    g = workloads.protein_grids(name)
    prot = cdk.Protein(ctx, g["values"], g["radii"], g["spacing"], g["mid"], g["length"])
    ctx.set_option("pipeline", 2)
    try:
        for seed, na, nq in ((911, 18, 5), (912, 45, 4), (913, 300, 2)):
            L = workloads.ligand(seed, na, 1)
            if na > 100:
                # compact coordinates: a ball of radius 6 A keeps the box well inside the grid
                rs = np.random.default_rng(seed)
                d = rs.normal(size=(na, 3))
                x = 6.0 * d / np.linalg.norm(d, axis=1, keepdims=True) * rs.uniform(0, 1, (na, 1)) ** (1 / 3)
                L["conformers"] = x[None].astype(np.float32)
            quats = workloads.quaternion_sets(seed + 1, 1, nq)
            lig = cdk.Ligand(prot, L["charge"], L["epsilon"], L["radius"])
            say("atom form", name, "ligand", na)
            E, idx, mc = cdk.search(ctx, prot, lig, L["conformers"], quats)
            assert ctx.get_option("last_pipeline") == 2
            prob = orc.PoseProblem(g["values"], g["radii"], g["spacing"], L["charge"], L["epsilon"], L["radius"])
            stats = check_search(prob, L["conformers"], quats, E, idx, mc)
            say("atom form", name, na, stats)
            assert stats.get("exact", 0) >= nq - 1
            for grp in (1, 3):
                ctx.set_option("mid_group", grp)
                E1, i1, _ = cdk.search(ctx, prot, lig, L["conformers"], quats)
                assert np.array_equal(E1.view(np.uint32), E.view(np.uint32)) and np.array_equal(i1, idx), grp
            ctx.set_option("mid_group", 0)
            lig.close()
    finally:
        ctx.set_option("mid_group", 0)
        ctx.set_option("pipeline", -1)
        prot.close()


@pytest.mark.parametrize("name", ["p96", "p128"])
def test_large_grids_vs_oracle(ctx, name):
    """96^3 and 128^3 grids, both engines, a few poses each."""
    g = workloads.protein_grids(name)
    prot = cdk.Protein(ctx, g["values"], g["radii"], g["spacing"], g["mid"], g["length"])
    L = workloads.ligand(77, 18, 1)
    quats = workloads.quaternion_sets(78, 1, 4)
    lig = cdk.Ligand(prot, L["charge"], L["epsilon"], L["radius"])
    prob = orc.PoseProblem(g["values"], g["radii"], g["spacing"], L["charge"], L["epsilon"], L["radius"])
    for engine in (cdk.ENGINE_REFERENCE, cdk.ENGINE_FUSED):
        say("grid", name, "engine", engine)
        E, idx, mc = cdk.search(ctx, prot, lig, L["conformers"], quats, engine=engine)
        stats = check_search(prob, L["conformers"], quats, E, idx, mc)
        say("grid", name, "engine", engine, stats)
        assert stats.get("exact", 0) >= 3
    lig.close()
    prot.close()
