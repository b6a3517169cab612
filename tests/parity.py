"""Parity rules shared by the GPU tests.

Tolerance (stated once, used everywhere):
  energy   |E_gpu - E64(t_gpu)| <= 1e-4 * max(|E64(t_gpu)|, S(t_gpu))
  argmin   t_gpu == t_oracle, or a documented near-tie:
           E64(t_gpu) - E64(t_oracle) <= 2e-4 * max(|E64(t_oracle)|, S(t_oracle))
where E64 is the oracle's FP64 dense correlation over the window and
S(t) = sum_c sum_r |L_c[r]| |P_c[r + t]| is the absolute-sum scale of that
translation (energies near zero lose relative accuracy in any FP32 FFT).
"""
import numpy as np

from oracle import oracle as orc

E_RTOL = 1e-4
TIE_RTOL = 2e-4


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
    stats["max_rel_err"] = max(stats.get("max_rel_err", 0.0), err / max(abs(E[g]), 1e-30))
    stats["max_scaled_err"] = max(stats.get("max_scaled_err", 0.0), err / max(abs(E[g]), Sg))
    if g != r:
        Sr = abs_scale(prob, L, r)
        gap = E[g] - E[r]
        assert gap <= TIE_RTOL * max(abs(E[r]), Sr), ("argmin", g, r, E[g], E[r], Sr)
        stats["near_ties"] = stats.get("near_ties", 0) + 1
    else:
        stats["exact"] = stats.get("exact", 0) + 1
    return o


def check_search(prob, conformers, quats, E, idx, mc, max_poses=None):
    stats = {}
    nconf, nquat = E.shape
    shared = quats.ndim == 2
    n = 0
    for c in range(nconf):
        for i in range(nquat):
            q = quats[i] if shared else quats[c, i]
            check_pose(prob, conformers[c], q, E[c, i], idx[c, i], mc[c, i], stats)
            n += 1
            if max_poses is not None and n >= max_poses:
                return stats
    return stats
