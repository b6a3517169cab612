"""Per-stage parity of the gfx950 kernels through the C-ABI.

Bit-exact where the reference's arithmetic is reproduced exactly (rotation,
AABB, FillLigandGrid spreading, argmin); within stated FP32 tolerances where
the arithmetic is a transform (FFTs, ConjMult, SumGrids), checked against
numpy's unnormalised rfftn/irfftn (the cuFFT R2C/C2R convention)."""
import os

import numpy as np
import pytest

import cdocker_amd as cdk
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = np.load(os.path.join(ROOT, "tests", "golden", "ref_vectors.npz"))
PROBE = np.array([1.2, 1.5, 1.7, 1.8, 1.9, 2.0, 2.1, 2.2], np.float32)


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def say(*a):
    print("[gpu]", *a, flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = cdk.Context(0)
    yield c
    c.close()


def test_rotate_bit_exact_golden(ctx):
    say("rotate golden")
    out = ctx.Rotate(GOLD["rot_quats"], GOLD["rot_coords"])
    assert np.array_equal(bits(out), bits(GOLD["rot_out"]))


def test_rotate_aabb_bit_exact(ctx):
    say("rotate_aabb golden")
    rot, mn, ln, win = ctx.rotate_aabb(GOLD["rot_quats"][:63], GOLD["rot_coords"], 64, 0.5)
    assert np.array_equal(bits(rot), bits(GOLD["rot_out"][:63]))
    assert np.array_equal(bits(mn), bits(GOLD["min_out"][:63]))
    assert np.array_equal(bits(ln), bits(GOLD["max_out"][:63] - GOLD["min_out"][:63]))
    expect = 64 - (ln / np.float32(0.5)).astype(np.int32) - 2
    assert np.array_equal(win, expect)


def test_fill_ligand_grid_bit_exact_golden(ctx):
    say("FillLigandGrid golden")
    nu = int(GOLD["fill_num_used"][0])
    ari = GOLD["fill_ari"]
    lists = [np.nonzero(ari == r)[0] for r in range(8)]
    dims = tuple(int(v) for v in GOLD["fill_dims"])
    g = ctx.FillLigandGrid(GOLD["fill_rot"], GOLD["fill_min"], GOLD["fill_charge"], GOLD["fill_eps"], nu,
                           GOLD["fill_used"], lists, *dims, float(GOLD["fill_spacing"][0]))
    flat = g.reshape(-1)
    nz = np.flatnonzero(flat)
    assert np.array_equal(nz, GOLD["fill_nz_idx"])
    assert np.array_equal(bits(flat[nz]), bits(GOLD["fill_nz_val"]))


def test_fill_accumulates_into_existing_values(ctx):
    say("FillLigandGrid +=")
    nu = int(GOLD["fill_num_used"][0])
    ari = GOLD["fill_ari"]
    lists = [np.nonzero(ari == r)[0] for r in range(8)]
    dims = tuple(int(v) for v in GOLD["fill_dims"])
    base = np.full(tuple(GOLD["fill_shape"]), 0.25, np.float32)
    g = ctx.FillLigandGrid(GOLD["fill_rot"], GOLD["fill_min"], GOLD["fill_charge"], GOLD["fill_eps"], nu,
                           GOLD["fill_used"], lists, *dims, float(GOLD["fill_spacing"][0]), grid=base.copy())
    o = orc.fill_ligand_grid(GOLD["fill_rot"], GOLD["fill_min"], GOLD["fill_charge"], GOLD["fill_eps"], nu,
                             GOLD["fill_used"], ari, 8, dims, float(GOLD["fill_spacing"][0]), grid=base.copy())
    assert np.array_equal(bits(g), bits(o))


@pytest.mark.parametrize("n", [32, 48, 64, 96, 128])
def test_fft_r2c_c2r_vs_numpy(ctx, n):
    say("fft n=%d" % n)
    rs = np.random.default_rng(n)
    b = 2
    x = rs.standard_normal((b, n, n, n)).astype(np.float32)
    X = ctx.fft_r2c_3d(x)
    ref = np.fft.rfftn(x.astype(np.float64), axes=(1, 2, 3))
    err = np.abs(X - ref).max() / np.abs(ref).max()
    assert err < 2e-6, err
    y = ctx.fft_c2r_3d(X)
    assert np.abs(y / n ** 3 - x).max() < 2e-5
    # C2R of an arbitrary (Hermitian-consistent) spectrum vs irfftn
    S = np.fft.rfftn(rs.standard_normal((1, n, n, n)), axes=(1, 2, 3)).astype(np.complex64)
    y = ctx.fft_c2r_3d(S)
    ref = np.fft.irfftn(S.astype(np.complex128), s=(n, n, n), axes=(1, 2, 3)) * n ** 3
    assert np.abs(y - ref).max() / np.abs(ref).max() < 2e-6


def test_fft_known_answers(ctx):
    say("fft known answers")
    n = 64
    x = np.zeros((1, n, n, n), np.float32)
    x[0, 0, 0, 0] = 1.0
    X = ctx.fft_r2c_3d(x)
    assert np.allclose(X, 1.0, atol=1e-6)  # impulse -> all ones
    x[:] = 0
    x[0, 3, 5, 7] = 1.0  # shift theorem
    X = ctx.fft_r2c_3d(x)
    k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n // 2 + 1), indexing="ij")
    ph = np.exp(-2j * np.pi * (3 * k[0] + 5 * k[1] + 7 * k[2]) / n)
    assert np.abs(X[0] - ph).max() < 2e-5
    # Parseval on the half spectrum: sum x^2 = (1/G) sum_k w_k |X_k|^2, w = 1
    # on the kz = 0 and kz = N/2 planes (their own mirrors) and 2 elsewhere
    x = np.random.default_rng(41).standard_normal((1, n, n, n)).astype(np.float32)
    X = ctx.fft_r2c_3d(x).astype(np.complex128)
    wk = np.full(n // 2 + 1, 2.0)
    wk[0] = wk[n // 2] = 1.0
    lhs = float((x.astype(np.float64) ** 2).sum())
    rhs = float((np.abs(X[0]) ** 2 * wk[None, None, :]).sum()) / n ** 3
    assert abs(lhs - rhs) <= 1e-5 * lhs, (lhs, rhs)


def test_conj_mult_and_sum_grids(ctx):
    say("ConjMult / SumGrids")
    rs = np.random.default_rng(4)
    nch, nq, H = 5, 3, 33 * 64 * 8
    P = (rs.standard_normal((nch, H)) + 1j * rs.standard_normal((nch, H))).astype(np.complex64)
    L = (rs.standard_normal((nq, nch, H)) + 1j * rs.standard_normal((nq, nch, H))).astype(np.complex64)
    out = ctx.ConjMult(P, L)
    ref = np.conj(L.astype(np.complex128)) * P.astype(np.complex128)[None]
    assert np.abs(out - ref).max() < 1e-5
    # conj placement: conj(L) * P, not L * conj(P)
    assert np.abs(out - L * np.conj(P)[None]).max() > 1e-2
    G = 64 ** 3
    S = ctx.SumGrids(out, nch, G)
    acc = np.zeros((nq, H), np.complex64)
    for c in range(nch):
        acc = (acc.real + out[:, c].real).astype(np.float32) + 1j * (acc.imag + out[:, c].imag).astype(np.float32)
    ref = (acc.real.astype(np.float32) / np.float32(np.sqrt(np.float32(G)))) + \
        1j * (acc.imag.astype(np.float32) / np.float32(np.sqrt(np.float32(G))))
    assert np.array_equal(S.real.view(np.uint32), ref.real.astype(np.float32).view(np.uint32))
    assert np.array_equal(S.imag.view(np.uint32), ref.imag.astype(np.float32).view(np.uint32))


def test_argmin_window_first_minimum(ctx):
    say("argmin")
    rs = np.random.default_rng(5)
    n = 32
    g = rs.standard_normal((4, n, n, n)).astype(np.float32)
    win = np.array([[30, 30, 30], [10, 20, 5], [1, 1, 1], [0, 5, 5]], np.int32)
    # planted ties: two equal minima, the lower (i, j, k) must win
    g[1, 3, 4, 2] = -50
    g[1, 2, 19, 4] = -50
    g[1, 7, 0, 0] = -50
    e, idx = ctx.argmin_window(g, win)
    sg = np.sqrt(np.float64(n ** 3))
    for q in range(3):
        E = (g[q, : win[q, 0], : win[q, 1], : win[q, 2]].astype(np.float64) / sg).astype(np.float32)
        em, im = orc.argmin_first(E)
        assert tuple(idx[q]) == tuple(im) and e[q] == np.float32(em)
    assert tuple(idx[1]) == (2, 19, 4)
    assert not np.isfinite(e[3]) and tuple(idx[3]) == (-1, -1, -1)  # empty window sentinel


def test_filter_quaternions_matches_reference(ctx):
    say("FilterQuaternions")
    out = ctx.FilterQuaternions(GOLD["filt_coords"], GOLD["filt_quats"], GOLD["filt_lens"], 10 ** 6)
    assert np.array_equal(bits(out), bits(GOLD["filt_out"]))
    capped = ctx.FilterQuaternions(GOLD["filt_coords"], GOLD["filt_quats"], GOLD["filt_lens"], 5, shuffle_seed=3)
    assert capped.shape[0] == 5
    assert all(any(np.array_equal(c, o) for o in GOLD["filt_out"]) for c in capped)
