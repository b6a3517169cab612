"""Grid rescoring and rigid refinement of retained poses on the GPU.

The score is pinned three ways: at poses the search returns it must equal the
oracle's FP64 per-atom trilinear gather (orc_sparse_energy_f64, written
independently from the search coordinates) and the search's own FP32 energy;
on hand-made grids it must reproduce closed forms (constant grids, linear
ramps); and its gradient must match central differences of the energy."""
import numpy as np
import pytest

import cdocker_amd as cdk
from cdocker_amd import workloads
from oracle import oracle as orc

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


def search_poses(ctx, prot, L, nq):
    lig = cdk.Ligand(prot, L["charge"], L["epsilon"], L["radius"])
    quats = L["quats"][0, :nq]
    E, idx, mc = cdk.search(ctx, prot, lig, L["conformers"][:1], quats)
    rot = ctx.Rotate(quats, L["conformers"][0])
    gmin = prot.grid_min()
    poses = rot.astype(np.float64) + (gmin[None, None, :] - mc[0][:, None, :].astype(np.float64)
                                      + (idx[0].astype(np.float32) * np.float32(prot.spacing))[:, None, :])
    return lig, quats, E[0], idx[0], mc[0], rot, poses


def test_rescore_equals_search_energy_and_oracle(ctx, p48):
    say("rescore vs search and oracle")
    g, prot = p48
    L = workloads.config1()["ligands"][0]
    lig, quats, E, idx, mc, rot, poses = search_poses(ctx, prot, L, 64)
    ok = np.isfinite(E)
    e, _, ins = cdk.rescore_poses(prot, lig, poses[ok])
    assert (ins == lig.natom).all()
    prob = orc.PoseProblem(g["values"], g["radii"], g["spacing"], L["charge"], L["epsilon"], L["radius"])
    e64 = np.array([prob.sparse_energy(rot[i], mc[i], idx[i]) for i in np.flatnonzero(ok)])
    scale = np.maximum(np.abs(e64), 1.0)
    assert np.abs(e - e64).max() / scale.max() < 1e-9
    assert (np.abs(e - E[ok].astype(np.float64)) <= 1e-4 * np.maximum(np.abs(e64), 10.0)).all()


def test_rescore_gradient_matches_central_differences(ctx, p48):
    say("rescore gradient")
    g, prot = p48
    L = workloads.config1()["ligands"][0]
    lig, quats, E, idx, mc, rot, poses = search_poses(ctx, prot, L, 8)
    rs = np.random.default_rng(3)
    x = poses[np.isfinite(E)][:4] + rs.uniform(-0.2, 0.2, size=(1, 1, 3))
    e0, grad, _ = cdk.rescore_poses(prot, lig, x, want_grad=True)
    h = 1e-6
    u = (x - prot.grid_min()[None, None, :]) / prot.spacing
    frac = u - np.floor(u)
    checked = 0
    for p in range(x.shape[0]):
        for a in range(lig.natom):
            for k in range(3):
                if not (0.01 < frac[p, a, k] < 0.99):
                    continue  # trilinear is piecewise linear; stay inside the cell
                xp = x[p:p + 1].copy()
                xm = x[p:p + 1].copy()
                xp[0, a, k] += h
                xm[0, a, k] -= h
                fd = (cdk.rescore_poses(prot, lig, xp)[0][0] - cdk.rescore_poses(prot, lig, xm)[0][0]) / (2 * h)
                assert abs(fd - grad[p, a, k]) <= 1e-5 * max(1.0, abs(grad[p, a, k])), (p, a, k, fd, grad[p, a, k])
                checked += 1
    assert checked > 100


def test_rescore_closed_forms(ctx):
    say("rescore closed forms")
    N, h = 32, 0.5
    radii = np.array([1.5, 2.0], np.float32)
    mid = np.array([8.0, 8.0, 8.0], np.float32)
    length = np.full(3, (N - 1) * h, np.float32)
    ix = np.arange(N, dtype=np.float32)
    vals = np.zeros((3, N, N, N), np.float32)
    vals[0] = 2.0                                   # constant vdW grid
    vals[1] = 0.25 * ix[:, None, None]              # ramp along x
    vals[2] = -0.5 * ix[None, None, :] + 1.0        # ramp along z (elec)
    prot = cdk.Protein(ctx, vals, radii, h, mid, length)
    q = np.array([0.3, -0.7, 0.4], np.float32)
    eps = np.array([0.16, 0.09, 0.25], np.float32)
    rad = np.array([1.5, 2.0, 1.5], np.float32)
    lig = cdk.Ligand(prot, q, eps, rad)
    gmin = prot.grid_min()
    x = gmin[None, :] + np.array([[3.3, 4.1, 5.7], [7.25, 2.5, 6.1], [10.0, 10.0, 1.05]])
    e, grad, ins = cdk.rescore_poses(prot, lig, x[None], want_grad=True)
    u = (x - gmin[None, :]) / h
    sv = np.sqrt(eps).astype(np.float64)  # sqrtf(eps) as the ligand tables hold it
    vdw = np.array([2.0, 0.25 * u[1, 0], 2.0])
    elec = -0.5 * u[:, 2] + 1.0
    expect = float((sv * vdw + q.astype(np.float64) * elec).sum())
    assert ins[0] == 3
    assert abs(e[0] - expect) < 1e-9
    gexp = np.zeros((3, 3))
    gexp[1, 0] = sv[1] * 0.25 / h
    gexp[:, 2] = q.astype(np.float64) * (-0.5) / h
    assert np.abs(grad[0] - gexp).max() < 1e-9
    # an atom off the grid contributes nothing
    far = x.copy()
    far[0] = gmin + 100.0
    e2, _, ins2 = cdk.rescore_poses(prot, lig, far[None])
    assert ins2[0] == 2
    assert abs(e2[0] - (expect - sv[0] * 2.0 - q[0] * elec[0])) < 1e-9
    lig.close()
    prot.close()


def test_refine_is_rigid_and_downhill(ctx, p48):
    say("refine")
    g, prot = p48
    L = workloads.config1()["ligands"][0]
    lig, quats, E, idx, mc, rot, poses = search_poses(ctx, prot, L, 16)
    x0 = poses[np.isfinite(E)][:8] + 0.3  # off the grid optimum
    e0, _, ins0 = cdk.rescore_poses(prot, lig, x0)
    e0 = e0 + 10.0 * (lig.natom - ins0)  # the refinement's off-grid penalty
    x1, e1 = cdk.refine_poses(prot, lig, x0, steps=60, off_grid_penalty=10.0)
    assert (e1 <= e0 + 1e-12).all() and (e1 < e0).any()
    d0 = np.linalg.norm(x0[:, :, None, :] - x0[:, None, :, :], axis=-1)
    d1 = np.linalg.norm(x1[:, :, None, :] - x1[:, None, :, :], axis=-1)
    assert np.abs(d0 - d1).max() < 1e-9
    er, _, ins = cdk.rescore_poses(prot, lig, x1)
    inside = ins == lig.natom
    assert np.allclose(er[inside], e1[inside], rtol=0, atol=1e-9)


def test_grid_for_openmm_order_and_bounds(ctx, p48):
    """CHARMM z-fastest grids to OpenMM's x-fastest order
    (AddGridForcesToOpenMMSystem.cpp:30-41) and the grid bounds in nm
    (TranRotaConfSearchGrid.cu:149-154)."""
    say("grid for openmm")
    g, prot = p48
    for c in (0, prot.num_grids - 1):
        v, b = cdk.grid_for_openmm(prot, c)
        assert np.array_equal(v, np.transpose(g["values"][c], (2, 1, 0)).astype(np.float64))
        gmin = prot.grid_min()
        gmax = gmin + np.float64(np.float32(g["n"] - 1) * np.float32(g["spacing"]))
        assert np.array_equal(b, np.concatenate([gmin * 0.1, gmax * 0.1]))


def test_protein_built_on_device_matches_oracle_and_upload(ctx):
    """cdk_protein_build: grids equal the oracle's brute-force builder bit for
    bit (same atom order, FP64); the spectra and a search equal those of the
    same grids uploaded from the host."""
    say("protein build")
    n, h, box = 32, 0.5, 16.0
    x, q, e, r = cdk.synth_protein_atoms(9, 400, box)
    v, radii, mid, ln = cdk.synth_protein(9, 400, box, n, h)
    built = cdk.Protein.build(ctx, x, q, e, r, radii, n, h, mid, ln)
    g = built.grids()
    gm = (np.float32(mid[0]) - np.float32(ln[0]) / np.float32(2)).astype(np.float64)
    ref = orc.build_grids(x, q, e, r, radii, n, np.full(3, gm), h)
    assert np.array_equal(g, ref)
    up = cdk.Protein(ctx, g, radii, h, mid, ln)
    for c in (0, 8):
        assert np.array_equal(built.spectrum(c).view(np.float32), up.spectrum(c).view(np.float32))
    L = workloads.ligand(31, 12, 1)
    quats = workloads.quaternion_sets(32, 1, 16)
    l1 = cdk.Ligand(built, L["charge"], L["epsilon"], L["radius"])
    l2 = cdk.Ligand(up, L["charge"], L["epsilon"], L["radius"])
    E1, i1, _ = cdk.search(ctx, built, l1, L["conformers"], quats)
    E2, i2, _ = cdk.search(ctx, up, l2, L["conformers"], quats)
    assert np.array_equal(E1.view(np.uint32), E2.view(np.uint32)) and np.array_equal(i1, i2)
    for o in (l1, l2, built, up):
        o.close()


def test_conformers_on_device_match_oracle(ctx):
    """Device random-torsion conformers: the kept trials and their order equal
    the oracle's; coordinates agree to 1e-9 A (atan2/sin/cos of the device
    and host math libraries may differ in the last bit)."""
    say("conformers")
    x, q, e, r = cdk.synth_ligand(21, 30)
    rot, mv, ex = orc.chain_rotors(30)
    conf, trial = cdk.generate_conformers(ctx, x, rot, mv, ex, 300, 5, 1.2, max_keep=50)
    ref, ok = orc.conformers(x.astype(np.float64), rot, mv, ex, 300, 5, 1.2)
    want = np.flatnonzero(ok)[:50]
    assert np.array_equal(trial, want)
    assert np.abs(conf - ref[want]).max() < 1e-9


def test_callers_at_bench_sizes(ctx):
    """The callers at the sizes the bench times them: the top 10 poses of each
    of 10 conformers of the 64^3 single-ligand search rescored (equal to the
    oracle's FP64 gather at the search's translation, and to the search's own
    energy), refined downhill, and 20,000 conformer trials of a 40-atom chain
    whose first 1,000 kept trials equal the oracle's."""
    g = workloads.protein_grids("p64")
    prot = cdk.Protein(ctx, g["values"], g["radii"], g["spacing"], g["mid"], g["length"])
    L = workloads.config2(nconf=10, nquat=200)["ligands"][0]
    lig = cdk.Ligand(prot, L["charge"], L["epsilon"], L["radius"])
    try:
        E, idx, mc = cdk.search(ctx, prot, lig, L["conformers"], L["quats"])
        prob = orc.PoseProblem(g["values"], g["radii"], g["spacing"], L["charge"], L["epsilon"], L["radius"])
        poses, e64, e32 = [], [], []
        for c in range(E.shape[0]):
            rot = ctx.Rotate(L["quats"][c], L["conformers"][c])
            rank, pose = cdk.rank_poses(prot, rot, E[c], idx[c], mc[c], 10)
            poses.append(pose)
            e64 += [prob.sparse_energy(rot[q], mc[c, q], idx[c, q]) for q in rank]
            e32 += [float(E[c, q]) for q in rank]
        poses = np.concatenate(poses)
        e64 = np.asarray(e64)
        say("callers: %d retained poses" % poses.shape[0])
        e, _, ins = cdk.rescore_poses(prot, lig, poses)
        assert poses.shape[0] == 100 and (ins == lig.natom).all()
        assert np.abs(e - e64).max() / np.maximum(np.abs(e64), 1.0).max() < 1e-9
        assert (np.abs(e - np.asarray(e32)) <= 1e-4 * np.maximum(np.abs(e64), 10.0)).all()
        x, e2 = cdk.refine_poses(prot, lig, poses, steps=50)
        assert (e2 <= e + 1e-9).all() and e2.mean() < e.mean()
        x0, _, _, _ = cdk.synth_ligand(4242, 40)
        rot, mv, ex = workloads.chain_topology(40)
        conf, trial = cdk.generate_conformers(ctx, x0, rot, mv, ex, 20000, 11, 1.2, max_keep=1000)
        ref, ok = orc.conformers(x0.astype(np.float64), rot, mv, ex, 20000, 11, 1.2)
        want = np.flatnonzero(ok)[:1000]
        say("callers: %d conformer trials kept of 20000" % int(ok.sum()))
        assert np.array_equal(trial, want) and np.abs(conf - ref[want]).max() < 1e-9
    finally:
        lig.close()
        prot.close()
