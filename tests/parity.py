"""Parity rules shared by the GPU tests.

Tolerance (stated once, used everywhere):
  energy   |E_gpu - E64(t_gpu)| <= 1e-4 * max(|E64(t_gpu)|, S(t_gpu)), and
           plain relative |E_gpu - E64(t_gpu)| <= 1e-4 * |E64(t_gpu)| wherever
           |E64(t_gpu)| >= 1e-2 * S(t_gpu)
  argmin   t_gpu == t_oracle, or a documented near-tie:
           E64(t_gpu) - E64(t_oracle) <= 2e-4 * max(|E64(t_oracle)|, S(t_oracle))
where E64 is the oracle's FP64 energy over the window and
S(t) = sum_c sum_r |L_c[r]| |P_c[r + t]| is the absolute-sum scale of that
translation (energies near zero lose relative accuracy in any FP32 FFT).
Every near-tie is appended to NEAR_TIES (pose, both translations, the FP64
gap and S); tests/conftest.py writes them to $CDK_NEAR_TIE_LOG at the end of
the session.
"""
import numpy as np

from oracle import oracle as orc

E_RTOL = 1e-4
TIE_RTOL = 2e-4
REL_FLOOR = 1e-2  # plain relative error asserted where |E64| >= REL_FLOOR * S
NEAR_TIES = []


def record_near_tie(label, pose, t_gpu, t_orc, e_gpu64, e_orc64, scale):
    NEAR_TIES.append(dict(test=label, pose=[int(v) for v in np.atleast_1d(pose)], t_gpu=[int(v) for v in t_gpu],
                          t_oracle=[int(v) for v in t_orc], e64_gpu=float(e_gpu64), e64_oracle=float(e_orc64),
                          gap=float(e_gpu64 - e_orc64), scale=float(scale),
                          gap_over_scale=float((e_gpu64 - e_orc64) / max(abs(e_orc64), scale))))


def abs_scale(prob, L, t):
    bx, by, bz = L.shape[1:]
    P = prob.Pused[:, t[0]:t[0] + bx, t[1]:t[1] + by, t[2]:t[2] + bz].astype(np.float64)
    return float((np.abs(L.astype(np.float64)) * np.abs(P)).sum())


def check_pose(prob, coords, q, e_gpu, idx_gpu, mc_gpu, stats):
    """Checks one pose; returns a dict describing it.  Raises on failure."""
    o = prob.search_pose(coords, q, double=True, want_E=True)
    assert np.array_equal(np.asarray(mc_gpu, np.float32).view(np.uint32), o["mincoor"].view(np.uint32)), \
        ("mincoor", mc_gpu, o["mincoor"])
    if not o["valid"]:
        assert not np.isfinite(e_gpu) and list(idx_gpu) == [-1, -1, -1], ("empty window", e_gpu, idx_gpu)
        stats["empty"] = stats.get("empty", 0) + 1
        return o
    g = tuple(int(v) for v in idx_gpu)
    r = tuple(int(v) for v in o["index"])
    E = o["E"]
    assert all(0 <= g[k] < E.shape[k] for k in range(3)), ("index outside window", g, E.shape)
    L, box, win, minc = prob.pose_box(coords, q)
    Sg = abs_scale(prob, L, g)
    err = abs(float(e_gpu) - E[g])
    assert err <= E_RTOL * max(abs(E[g]), Sg), ("energy", float(e_gpu), E[g], Sg)
    if abs(E[g]) >= REL_FLOOR * Sg:
        assert err <= E_RTOL * abs(E[g]), ("relative energy", float(e_gpu), E[g], Sg)
    stats["max_rel_err"] = max(stats.get("max_rel_err", 0.0), err / max(abs(E[g]), 1e-30))
    stats["max_scaled_err"] = max(stats.get("max_scaled_err", 0.0), err / max(abs(E[g]), Sg))
    if g != r:
        Sr = abs_scale(prob, L, r)
        gap = E[g] - E[r]
        assert gap <= TIE_RTOL * max(abs(E[r]), Sr), ("argmin", g, r, E[g], E[r], Sr)
        stats["near_ties"] = stats.get("near_ties", 0) + 1
        record_near_tie(stats.get("label", ""), stats.get("n", 0), g, r, E[g], E[r], Sr)
    else:
        stats["exact"] = stats.get("exact", 0) + 1
    return o


def check_sampled(prob, conformers, conf, quats, E, idx, mc, label, argmin=True):
    """Checks many poses against the oracle's FP64 per-atom gather
    (orc_sparse_search_f64): pose p is conformer conformers[conf[p]] rotated
    by quats[p] with GPU results E[p], idx[p], mc[p].  Mincoor bit for bit;
    energy at the GPU's translation and, with argmin, the translation itself
    over the whole window, under the rules above.  Returns stats with the
    largest scaled and plain relative errors."""
    confs = np.asarray(conformers, np.float32)
    conf = np.asarray(conf, np.int32)
    quats = np.asarray(quats, np.float32).reshape(-1, 4)
    E = np.asarray(E, np.float32).reshape(-1)
    idx = np.asarray(idx, np.int32).reshape(-1, 3)
    mc = np.asarray(mc, np.float32).reshape(-1, 3)
    n = quats.shape[0]
    for c in np.unique(conf):
        sel = np.flatnonzero(conf == c)
        rot = orc.rotate(quats[sel], confs[c])
        mn, _ = orc.min_max_coors(rot)
        assert np.array_equal(mn.view(np.uint32), mc[sel].view(np.uint32)), ("mincoor", label, int(c))
    o = prob.sparse_search(confs, conf, quats, t_at=np.where(idx < 0, 0, idx), full=argmin)
    stats = dict(poses=n, valid=int(o["valid"].sum()), exact=0, near_ties=0, max_scaled_err=0.0, max_rel_err=0.0,
                 max_rel_err_where_asserted=0.0)
    for p in range(n):
        if not o["valid"][p]:
            assert not np.isfinite(E[p]) and list(idx[p]) == [-1, -1, -1], ("empty window", label, p, E[p], idx[p])
            continue
        assert np.isfinite(E[p]) and np.isfinite(o["e_at"][p]), ("index outside window", label, p, idx[p])
        e64, S = o["e_at"][p], o["s_at"][p]
        err = abs(float(E[p]) - e64)
        assert err <= E_RTOL * max(abs(e64), S), ("energy", label, p, float(E[p]), e64, S)
        rel = err / max(abs(e64), 1e-30)
        if abs(e64) >= REL_FLOOR * S:
            assert rel <= E_RTOL, ("relative energy", label, p, float(E[p]), e64, S)
            stats["max_rel_err_where_asserted"] = max(stats["max_rel_err_where_asserted"], rel)
        stats["max_rel_err"] = max(stats["max_rel_err"], rel)
        stats["max_scaled_err"] = max(stats["max_scaled_err"], err / max(abs(e64), S))
        if argmin:
            t_orc = o["index"][p]
            if tuple(idx[p]) == tuple(t_orc):
                stats["exact"] += 1
            else:
                Sr = prob.sparse_search(confs, conf[p:p + 1], quats[p:p + 1], t_at=t_orc[None], full=False)["s_at"][0]
                gap = e64 - o["emin"][p]
                assert gap <= TIE_RTOL * max(abs(o["emin"][p]), Sr), ("argmin", label, p, idx[p], t_orc, gap, Sr)
                stats["near_ties"] += 1
                record_near_tie(label, p, idx[p], t_orc, e64, o["emin"][p], Sr)
    return stats


def check_search(prob, conformers, quats, E, idx, mc, max_poses=None, label=""):
    stats = {"label": label}
    nconf, nquat = E.shape
    shared = quats.ndim == 2
    n = 0
    for c in range(nconf):
        for i in range(nquat):
            q = quats[i] if shared else quats[c, i]
            stats["n"] = [c, i]
            check_pose(prob, conformers[c], q, E[c, i], idx[c, i], mc[c, i], stats)
            n += 1
            if max_poses is not None and n >= max_poses:
                break
        if max_poses is not None and n >= max_poses:
            break
    stats.pop("label", None)
    stats.pop("n", None)
    return stats
