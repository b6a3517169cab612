"""Pins the oracle (CPU restatement) before anything is checked against it.

* golden vectors written by the reference's own compiled sources
  (tests/golden/ref_vectors.npz): bit-exact;
* the compiled reference subset live, on fresh inputs, when it is built;
* closed forms and invariants (rotation matrix, spreading mass and first
  moment, constant-grid, linear-ramp and single-atom correlations);
* an independent FP64 formulation (per-atom trilinear gather) and an FFT
  restatement of the reference's correlation in numpy;
* planted faults that those checks must catch.
"""
import os

import numpy as np
import pytest

from oracle import oracle as orc

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz"))
PROBE = np.array([1.2, 1.5, 1.7, 1.8, 1.9, 2.0, 2.1, 2.2], np.float32)


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


# ---------------------------------------------------------------- golden
def test_rotate_golden_bit_exact():
    out = orc.rotate(GOLD["rot_quats"], GOLD["rot_coords"])
    assert np.array_equal(bits(out), bits(GOLD["rot_out"]))


def test_min_max_golden_bit_exact():
    mn, mx = orc.min_max_coors(GOLD["rot_out"])
    assert np.array_equal(bits(mn), bits(GOLD["min_out"]))
    assert np.array_equal(bits(mx), bits(GOLD["max_out"]))
    tmn, tmx = orc.min_max_coors(GOLD["tie_coords"])
    assert np.array_equal(bits(tmn), bits(GOLD["tie_min"]))  # signed zeros: first occurrence wins
    assert np.array_equal(bits(tmx), bits(GOLD["tie_max"]))


def test_idx_of_atoms_golden():
    nu, used, ari = orc.idx_of_atoms_for_vdw_radius(GOLD["idx_radii"], PROBE)
    assert nu == int(GOLD["idx_num_used"][0])
    assert np.array_equal(used, GOLD["idx_used"])
    assert np.array_equal(ari, GOLD["idx_ari"])


def test_fill_ligand_grid_golden_bit_exact():
    g = orc.fill_ligand_grid(GOLD["fill_rot"], GOLD["fill_min"], GOLD["fill_charge"], GOLD["fill_eps"],
                             int(GOLD["fill_num_used"][0]), GOLD["fill_used"], GOLD["fill_ari"], 8,
                             tuple(GOLD["fill_dims"]), float(GOLD["fill_spacing"][0]))
    assert tuple(g.shape) == tuple(GOLD["fill_shape"])
    flat = g.reshape(-1)
    nz = np.flatnonzero(flat)
    assert np.array_equal(nz, GOLD["fill_nz_idx"])
    assert np.array_equal(bits(flat[nz]), bits(GOLD["fill_nz_val"]))


def test_quaternion_sampling_golden_bit_exact():
    out = orc.quaternion_uniform_sampling(int(GOLD["qus_seed"][0]), 3, 40)
    assert np.array_equal(bits(out), bits(GOLD["qus_out"]))


def test_filter_golden():
    out = orc.filter_quaternions(GOLD["filt_coords"], GOLD["filt_quats"], GOLD["filt_lens"])
    assert np.array_equal(bits(out), bits(GOLD["filt_out"]))
    assert 20 < out.shape[0] < GOLD["filt_quats"].shape[0]


# ------------------------------------------------- live compiled reference
needs_ref = pytest.mark.skipif(not orc.have_ref(), reason="oracle/_ref not built (reference tree absent)")


@needs_ref
def test_live_reference_fresh_inputs():
    rs = np.random.default_rng(77)
    for trial in range(5):
        n = int(rs.integers(5, 60))
        x = (rs.standard_normal((n, 3)) * 4).astype(np.float32)
        q = rs.standard_normal((16, 4)).astype(np.float32)
        q /= np.linalg.norm(q, axis=1, keepdims=True)
        a = orc.rotate(q, x)
        b = orc.rotate(q, x, lib="ref")
        assert np.array_equal(bits(a), bits(b))
        mna, mxa = orc.min_max_coors(a)
        mnb, mxb = orc.min_max_coors(b, lib="ref")
        assert np.array_equal(bits(mna), bits(mnb)) and np.array_equal(bits(mxa), bits(mxb))
        r = rs.choice(PROBE, n).astype(np.float32)
        assert all(np.array_equal(u, v) for u, v in zip(orc.idx_of_atoms_for_vdw_radius(r, PROBE)[1:],
                                                        orc.idx_of_atoms_for_vdw_radius(r, PROBE, lib="ref")[1:]))
        nu, used, ari = orc.idx_of_atoms_for_vdw_radius(r, PROBE)
        ch = (0.3 * rs.standard_normal(n)).astype(np.float32)
        ep = rs.uniform(0.05, 0.2, n).astype(np.float32)
        dims = (40, 40, 40)
        g1 = orc.fill_ligand_grid(a[:4], mna[:4], ch, ep, nu, used, ari, 8, dims, 0.5)
        g2 = orc.fill_ligand_grid(a[:4], mna[:4], ch, ep, nu, used, ari, 8, dims, 0.5, lib="ref")
        assert np.array_equal(bits(g1), bits(g2))
    s1 = orc.quaternion_uniform_sampling(5, 4, 33)
    s2 = orc.quaternion_uniform_sampling(5, 4, 33, lib="ref")
    assert np.array_equal(bits(s1), bits(s2))


@needs_ref
def test_live_reference_grid_reader_reproduces_fixture():
    """The grid-reader fixture is what the compiled ReadGrids.cpp reads."""
    here = os.path.join(os.path.dirname(__file__), "golden")
    gold = np.load(os.path.join(here, "grids_ref.npz"))
    for name in ("plain", "crlf"):
        v, radii, mid, ln, sp, rk = orc.read_grids_ref(os.path.join(here, "grids_%s.txt" % name))
        assert np.array_equal(bits(v), bits(gold[name + "_values"]))
        assert np.array_equal(bits(radii), bits(gold[name + "_radii"]))
        assert np.array_equal(mid, gold[name + "_mid"]) and sp == gold[name + "_spacing"][0]


# ------------------------------------------------------ closed forms
def rotmat(q):
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def test_rotate_closed_form():
    rs = np.random.default_rng(1)
    x = (rs.standard_normal((30, 3)) * 5).astype(np.float32)
    q = rs.standard_normal((50, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    out = orc.rotate(q.astype(np.float32), x)
    for i in range(50):
        ref = x.astype(np.float64) @ rotmat(q[i]).T
        assert np.allclose(out[i], ref, atol=5e-5)
    # identity and q == -q
    assert np.array_equal(orc.rotate(np.array([1, 0, 0, 0], np.float32), x)[0], x)
    assert np.array_equal(bits(orc.rotate(q[:5].astype(np.float32), x)),
                          bits(orc.rotate(-q[:5].astype(np.float32), x)))


def test_spread_mass_and_first_moment():
    rs = np.random.default_rng(2)
    n = 25
    x = (rs.standard_normal((n, 3)) * 3).astype(np.float32)
    r = rs.choice(PROBE, n).astype(np.float32)
    ch = rs.standard_normal(n).astype(np.float32)
    ep = rs.uniform(0.05, 0.2, n).astype(np.float32)
    nu, used, ari = orc.idx_of_atoms_for_vdw_radius(r, PROBE)
    q = np.array([[0.8, 0.2, -0.4, 0.4]], np.float32)
    q /= np.linalg.norm(q)
    rot = orc.rotate(q, x)
    mn, _ = orc.min_max_coors(rot)
    h = 0.5
    g = orc.fill_ligand_grid(rot, mn, ch, ep, nu, used, ari, 8, (40, 40, 40), h)[0]
    idx = np.indices((40, 40, 40)).reshape(3, -1).T
    for c in range(nu + 1):
        atoms = np.nonzero(ari == used[c])[0] if c < nu else np.arange(n)
        w = np.sqrt(ep[atoms]).astype(np.float64) if c < nu else ch[atoms].astype(np.float64)
        assert abs(g[c].sum(dtype=np.float64) - w.sum()) < 1e-5 * (1 + np.abs(w).sum())
        mom = (g[c].reshape(-1, 1).astype(np.float64) * idx).sum(0)
        expect = (w[:, None] * (rot[0][atoms].astype(np.float64) - mn[0]) / h).sum(0)
        assert np.allclose(mom, expect, atol=1e-4 * (1 + np.abs(expect).max()))


def test_single_atom_on_node():
    x = np.array([[0.0, 0.0, 0.0]], np.float32)
    g = orc.fill_ligand_grid(x[None], np.zeros((1, 3), np.float32), np.array([0.7], np.float32),
                             np.array([0.09], np.float32), 1, np.array([3]), np.array([3]), 8, (6, 6, 6), 0.5)[0]
    assert g[1, 0, 0, 0] == np.float32(0.7) and np.count_nonzero(g[1]) == 1
    assert g[0, 0, 0, 0] == np.float32(0.3) and np.count_nonzero(g[0]) == 1


# --------------------------------------------- correlation (cpuNaive)
def small_problem(seed=3, N=32, natom=12, h=0.5):
    rs = np.random.default_rng(seed)
    vals = rs.standard_normal((9, N, N, N)).astype(np.float32)
    x = (rs.standard_normal((natom, 3)) * 1.6).astype(np.float32)
    ch = (0.3 * rs.standard_normal(natom)).astype(np.float32)
    ep = rs.uniform(0.05, 0.2, natom).astype(np.float32)
    r = rs.choice(PROBE, natom).astype(np.float32)
    return orc.PoseProblem(vals, PROBE, h, ch, ep, r), x, rs


def test_dense_matches_sparse_fp64():
    prob, x, rs = small_problem()
    q = rs.standard_normal(4).astype(np.float32)
    q /= np.linalg.norm(q)
    res = prob.search_pose(x, q, double=True, want_E=True)
    assert res["valid"]
    rot = orc.rotate(q, x)[0]
    E = res["E"]
    # The dense form correlates the FP32-rounded ligand cells (as the
    # reference does); the sparse form uses unrounded weights, so they agree
    # to float rounding of the cells.  A dropped atom or corner is >= 1e-2.
    for t in [(0, 0, 0), tuple(np.array(E.shape) - 1), tuple(res["index"]), (3, 1, 2)]:
        s = prob.sparse_energy(rot, res["mincoor"], t)
        assert abs(E[t] - s) <= 2e-6 * (1 + abs(s))
    # planted fault: dropping one atom's contribution must be detected
    prob.charge = prob.charge.copy()
    prob.charge[0] += np.float32(0.05)
    s_bad = prob.sparse_energy(rot, res["mincoor"], tuple(res["index"]))
    assert abs(E[tuple(res["index"])] - s_bad) > 1e-3


def test_sparse_search_matches_dense_fp64_search():
    """orc_sparse_search_f64 (the full-window gather search used to check
    argmins at scale) against the dense FP64 correlation's search of the same
    poses: same validity, same argmin (or a gap within the cells' float
    rounding), same minimum; E64 and S at a given translation equal the
    single-translation gather."""
    prob, x, rs = small_problem(seed=9)
    confs = np.stack([x, x[::-1].copy()])  # two conformers (the second a relabelling)
    q = rs.standard_normal((12, 4)).astype(np.float32)
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    conf = np.arange(12) % 2
    t_at = np.zeros((12, 3), np.int32)
    t_at[:, 0] = 1
    t_at[:, 2] = 2
    res = prob.sparse_search(confs, conf, q, t_at)
    nvalid = 0
    for p in range(12):
        d = prob.search_pose(confs[conf[p]], q[p], double=True, want_E=True)
        assert bool(res["valid"][p]) == d["valid"]
        if not d["valid"]:
            continue
        nvalid += 1
        E = d["E"]
        gi = tuple(res["index"][p])
        assert abs(res["emin"][p] - E.min()) <= 2e-6 * (1 + abs(E.min()))
        assert E[gi] - E.min() <= 4e-6 * (1 + abs(E.min()))
        rot = orc.rotate(q[p], confs[conf[p]])[0]
        e1 = prob.sparse_energy(rot, d["mincoor"], (1, 0, 2))
        assert abs(res["e_at"][p] - e1) <= 1e-12 * (1 + abs(e1))  # same terms, other summation order
        assert res["s_at"][p] >= abs(res["e_at"][p])
    assert nvalid >= 8
    # planted fault: a transposed translation (k, j, i) is detected
    p = int(np.flatnonzero(res["valid"])[0])
    d = prob.search_pose(confs[conf[p]], q[p], double=True, want_E=True)
    i, j, k = res["index"][p]
    if (i, j, k) != (k, j, i) and max(i, k) < min(d["E"].shape[0], d["E"].shape[2]):
        assert d["E"][k, j, i] - d["E"].min() > 1e-6


def test_timed_cpu_baseline_build_rounds_as_the_oracle():
    """libcpubase.so (the timed CPU baseline: the same source at -O3 for
    AVX2/FMA hosts, contraction off) gives the oracle's FP32 cpuNaive loop
    bit for bit."""
    prob, x, rs = small_problem(seed=6)
    q = rs.standard_normal(4).astype(np.float32)
    q /= np.linalg.norm(q)
    L, box, win, minc = prob.pose_box(x, q)
    a = orc.correlate_window_full(prob.Pused, L, win)
    b = orc.correlate_window_full_timed(prob.Pused, L, win, 3)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_dense_matches_fft_restatement():
    """The reference computes IFFT(sum_c conj(FFT L_c) FFT P_c)/G; in FP64 it
    must equal the direct correlation on the window."""
    prob, x, rs = small_problem(seed=4)
    q = rs.standard_normal(4).astype(np.float32)
    q /= np.linalg.norm(q)
    L, box, win, minc = prob.pose_box(x, q)
    N = prob.N
    Lf = np.zeros((L.shape[0], N, N, N))
    Lf[:, : box[0], : box[1], : box[2]] = L
    S = (np.conj(np.fft.rfftn(Lf, axes=(1, 2, 3))) * np.fft.rfftn(prob.Pused.astype(np.float64), axes=(1, 2, 3)))
    Efft = np.fft.irfftn(S.sum(0), s=(N, N, N), axes=(0, 1, 2))
    Ed = orc.correlate_window(prob.Pused, L, win, double=True)
    assert np.allclose(Efft[: win[0], : win[1], : win[2]], Ed, rtol=1e-9, atol=1e-9)
    # planted fault: transposing the ligand box (x<->z) must be detected
    Lt = np.ascontiguousarray(np.transpose(L, (0, 3, 2, 1)))
    bad = orc.correlate_window(prob.Pused, Lt, (win[2], win[1], win[0]), double=True)
    assert not np.allclose(bad[: min(win), : min(win), : min(win)], Ed[: min(win), : min(win), : min(win)],
                           rtol=1e-3)


def test_dense_fp32_is_cpunaive_order():
    """FP32 accumulation in c, m, n, o order: a plain-loop restatement on a
    tiny box must agree bit for bit (zero cells included)."""
    rs = np.random.default_rng(5)
    N = 10
    P = rs.standard_normal((2, N, N, N)).astype(np.float32)
    L = rs.standard_normal((2, 3, 3, 4)).astype(np.float32)
    L[L < 0.3] = 0
    win = (N - 3, N - 3, N - 4)
    E = orc.correlate_window(P, L, win)
    ref = np.zeros(win, np.float32)
    for i in range(win[0]):
        for j in range(win[1]):
            for k in range(win[2]):
                e = np.float32(0)
                for c in range(2):
                    for m in range(3):
                        for n in range(3):
                            for o in range(4):
                                e = np.float32(e + np.float32(L[c, m, n, o] * P[c, m + i, n + j, o + k]))
                ref[i, j, k] = e
    assert np.array_equal(bits(E), bits(ref))
    # the literal loop (zeros included, the timed CPU baseline) is bit-identical
    assert np.array_equal(bits(orc.correlate_window_full(P, L, win)), bits(ref))


def test_constant_grid_known_answer():
    prob, x, rs = small_problem(seed=6)
    k = np.arange(1, prob.Pused.shape[0] + 1, dtype=np.float32) * np.float32(0.25)
    prob.Pused = np.ascontiguousarray(np.broadcast_to(k[:, None, None, None], prob.Pused.shape).astype(np.float32))
    q = np.array([1, 0, 0, 0], np.float32)
    res = prob.search_pose(x, q, double=True, want_E=True)
    expect = 0.0
    for c in range(prob.num_used):
        expect += k[c] * np.sqrt(prob.eps[prob.ari == prob.idx_used[c]]).astype(np.float64).sum()
    expect += k[-1] * prob.charge.astype(np.float64).sum()
    assert np.allclose(res["E"], expect, rtol=1e-6, atol=1e-6)


def test_linear_ramp_grids_closed_form():
    """Every protein channel a different linear ramp P_c(i, j, k) = a_c i +
    b_c j + g_c k + d_c (dyadic coefficients: exact in FP32).  Trilinear
    spreading reproduces a linear function exactly, so the FP64 correlation
    at translation t is sum_c sum_{a in c} w_a P_c(u_a + t) with u_a = (p_a -
    m) / h the atom's continuous grid position: linear in t with slopes
    sum_c (a_c, b_c, g_c) W_c.  A transposed axis, a wrong sign of t or a
    dropped channel each change it."""
    prob, x, rs = small_problem(seed=11)
    N = prob.N
    C = prob.Pused.shape[0]
    coef = np.array([[0.25, -0.5, 0.125, 1.0], [-0.125, 0.25, 0.5, -2.0], [0.5, 0.125, -0.25, 0.5],
                     [0.0625, -0.25, 0.375, 0.0], [-0.375, 0.5, 0.0625, 3.0], [0.125, 0.125, -0.5, -1.0],
                     [0.25, 0.375, 0.25, 0.25], [-0.5, -0.0625, 0.125, 2.0], [0.375, -0.125, -0.375, -0.5]])[:C]
    i, j, k = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    prob.Pused = np.ascontiguousarray(np.stack(
        [(a * i + b * j + g * k + d).astype(np.float32) for a, b, g, d in coef]))
    q = rs.standard_normal(4).astype(np.float32)
    q /= np.linalg.norm(q)
    res = prob.search_pose(x, q, double=True, want_E=True)
    assert res["valid"]
    rot = orc.rotate(q, x)[0].astype(np.float64)
    u = (rot - res["mincoor"].astype(np.float64)) / prob.h
    ti, tj, tk = np.meshgrid(*[np.arange(n) for n in res["window"]], indexing="ij")

    def closed_form(axes):
        """sum over channels of the ramp at the atoms' positions plus t; axes
        permutes which coefficient goes with which grid axis."""
        E = np.zeros(tuple(res["window"]))
        for c in range(C):
            elec = c == C - 1
            sel = np.ones(prob.natom, bool) if elec else prob.atom_channel == c
            w = prob.charge.astype(np.float64) if elec else np.sqrt(prob.eps.astype(np.float64))
            W = w[sel].sum()
            Wu = (w[sel, None] * u[sel]).sum(axis=0)
            s = coef[c][list(axes)]
            E += s[0] * (Wu[0] + ti * W) + s[1] * (Wu[1] + tj * W) + s[2] * (Wu[2] + tk * W) + coef[c][3] * W
        return E

    E = closed_form((0, 1, 2))
    scale = np.abs(E).max()
    assert np.abs(res["E"] - E).max() <= 1e-6 * scale, np.abs(res["E"] - E).max() / scale
    # the x and z axes swapped (a transposed operand) does not fit
    assert np.abs(res["E"] - closed_form((2, 1, 0))).max() > 1e-3 * scale


def test_single_atom_correlation_is_window_of_grid():
    rs = np.random.default_rng(8)
    N = 32
    vals = rs.standard_normal((9, N, N, N)).astype(np.float32)
    prob = orc.PoseProblem(vals, PROBE, 0.5, np.array([0.5], np.float32), np.array([0.16], np.float32),
                           np.array([1.8], np.float32))
    res = prob.search_pose(np.zeros((1, 3), np.float32), np.array([1, 0, 0, 0], np.float32), double=True,
                           want_E=True)
    w = res["window"]
    expect = 0.4 * vals[3][: w[0], : w[1], : w[2]].astype(np.float64) + 0.5 * vals[8][: w[0], : w[1], : w[2]]
    assert np.allclose(res["E"], expect, rtol=1e-6, atol=1e-6)
    # and the argmin is the textbook argmin of that window
    assert tuple(res["index"]) == np.unravel_index(np.argmin(expect), expect.shape)


def test_argmin_first_minimum_wins():
    E = np.ones((4, 5, 6), np.float32)
    E[2, 3, 1] = -1
    E[1, 4, 5] = -1
    E[3, 0, 0] = -1
    e, idx = orc.argmin_first(E)
    assert e == -1 and tuple(idx) == (1, 4, 5)
    E[0, 0, 0] = np.nan
    e, idx = orc.argmin_first(E)
    assert tuple(idx) == (1, 4, 5)


def test_window_and_empty_window():
    prob, x, rs = small_problem(seed=9, N=32)
    big = x * 8  # extent beyond the grid -> empty window, sentinel
    res = prob.search_pose(big, np.array([1, 0, 0, 0], np.float32))
    assert not res["valid"] and not np.isfinite(res["energy"]) and list(res["index"]) == [-1, -1, -1]
    res = prob.search_pose(x, np.array([1, 0, 0, 0], np.float32))
    ln = np.ptp(x, axis=0).astype(np.float32)
    assert list(res["window"]) == [32 - int(np.float32(l) / np.float32(0.5)) - 2 for l in ln]


def test_grid_builder_oracle_closed_forms():
    """orc_build_grids: one atom reproduces the LJ / Coulomb formulas at every
    grid point (cutoffs and caps included); two atoms add; a planted charge
    change is seen."""
    n, h = 32, 0.5
    gmin = np.array([-3.0, -2.5, -4.0])
    probes = np.array([1.2, 1.7, 2.2], np.float32)
    xyz = np.array([[1.1, 2.3, 0.7]])
    q, eps, rad = np.array([0.4]), np.array([0.16]), np.array([1.5])
    g = orc.build_grids(xyz, q, eps, rad, probes, n, gmin, h)
    ax = gmin[:, None] + np.arange(n)[None, :] * h
    d2 = ((ax[0][:, None, None] - xyz[0, 0]) ** 2 + (ax[1][None, :, None] - xyz[0, 1]) ** 2
          + (ax[2][None, None, :] - xyz[0, 2]) ** 2)
    d2c = np.maximum(d2, 1e-4)
    el = np.where(d2 <= 144.0, 332.0636 * q[0] / (4.0 * d2c), 0.0)
    assert np.allclose(g[3], np.clip(el, -20, 20).astype(np.float32), rtol=1e-6, atol=1e-6)
    for p in range(3):
        s6 = ((rad[0] + np.float64(probes[p])) ** 2 / d2c) ** 3
        lj = np.where(d2 <= 64.0, np.sqrt(eps[0]) * (s6 * s6 - 2 * s6), 0.0)
        assert np.allclose(g[p], np.minimum(lj, 3.0).astype(np.float32), rtol=1e-6, atol=1e-6)
    assert (g[:3] == 3.0).any() and (g[3] == 20.0).any()  # caps reached near the atom
    assert (g[3] == 0.0).any() and (g[0] == 0.0).any()    # cutoffs reached far away
    # a second, distant atom adds its own field (no cap involved far from both)
    xyz2 = np.array([[1.1, 2.3, 0.7], [9.0, 9.0, 9.0]])
    g2 = orc.build_grids(xyz2, np.array([0.4, -0.3]), np.array([0.16, 0.09]), np.array([1.5, 1.8]), probes, n,
                         gmin, h)
    gb = orc.build_grids(xyz2[1:], np.array([-0.3]), np.array([0.09]), np.array([1.8]), probes, n, gmin, h)
    far = (np.abs(g) < 1.0) & (np.abs(gb) < 1.0)
    assert np.allclose(g2[far], (g.astype(np.float64) + gb)[far], rtol=1e-5, atol=1e-6)
    assert not np.array_equal(orc.build_grids(xyz, q + 0.01, eps, rad, probes, n, gmin, h), g)


def _dihedral(p0, p1, p2, p3):
    b0, b1, b2 = p0 - p1, p2 - p1, p3 - p2
    b1n = b1 / np.linalg.norm(b1)
    v = b0 - np.dot(b0, b1n) * b1n
    w = b2 - np.dot(b2, b1n) * b1n
    return np.arctan2(np.dot(np.cross(b1n, v), w), np.dot(v, w))


def _torsion_u(seed, t, r):
    m = (1 << 64) - 1
    z = (seed + 0x9E3779B97F4A7C15 * ((t * 0x100000001 + r + 1) & m)) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return ((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0)


def test_conformer_oracle_dihedrals_rigidity_and_filter():
    """orc_conformers: every rotor ends at its drawn IUPAC dihedral (a numpy
    dihedral, independent of the oracle's formula), bond lengths and angles
    are unchanged, and the keep flags equal a brute-force minimum-distance
    check; another seed gives other conformers."""
    rs = np.random.default_rng(8)
    n = 18
    x = np.zeros((n, 3))
    for i in range(1, n):  # a bent chain, 1.5 A bonds
        v = rs.normal(size=3)
        x[i] = x[i - 1] + 1.5 * v / np.linalg.norm(v)
    rot, mv, ex = orc.chain_rotors(n)
    out, ok = orc.conformers(x, rot, mv, ex, 40, 11, 1.3)
    for t in range(40):
        for r, (a, b, c, d) in enumerate(rot):
            th = -np.pi + 2 * np.pi * _torsion_u(11, t, r)
            err = np.angle(np.exp(1j * (_dihedral(*out[t, [a, b, c, d]]) - th)))
            assert abs(err) < 1e-9
        bl0 = np.linalg.norm(np.diff(x, axis=0), axis=1)
        bl = np.linalg.norm(np.diff(out[t], axis=0), axis=1)
        assert np.abs(bl - bl0).max() < 1e-12
        d2 = ((out[t][:, None] - out[t][None]) ** 2).sum(-1)
        iu = np.triu_indices(n, 1)
        keep = not ((d2[iu] < 1.3 ** 2) & (ex[iu] == 0)).any()
        assert keep == ok[t]
    assert 0 < ok.sum() < 40
    out2, _ = orc.conformers(x, rot, mv, ex, 40, 12, 1.3)
    assert not np.allclose(out, out2)
