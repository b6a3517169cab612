"""Writes tests/golden/ref_vectors.npz from the reference's own pure-math
sources (Rotate.cpp, QuaternionMultiply.cpp, GetMinCoors.cpp, GetMaxCoors.cpp,
GetIdxOfAtomsForVdwRadius.cpp, FillLigandGrid.cpp,
QuaternionUniformSampling.cpp, FilterQuaternions.cpp), compiled by
`make -C oracle ref` into oracle/_ref/libcdkref.so; and grid text files
tests/golden/grids_*.txt (written here, in the ReadGrids format with assorted
spacing, line endings and number spellings) with what the reference's
ReadGrids.cpp reads from them in tests/golden/grids_ref.npz.

Inputs are drawn with numpy's seeded generator, independently of the
library's own generators.  Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as orc  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "ref_vectors.npz")
GRID_OUT = os.path.join(HERE, "grids_ref.npz")
PROBE = np.array([1.2, 1.5, 1.7, 1.8, 1.9, 2.0, 2.1, 2.2], np.float32)


def unit_quats(rs, n):
    q = rs.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return q.astype(np.float32)


def walk(rs, n, step=1.5):
    p = np.zeros((n, 3))
    for i in range(1, n):
        while True:
            v = rs.standard_normal(3)
            c = p[i - 1] + step * v / np.linalg.norm(v)
            if i < 2 or np.min(np.linalg.norm(p[: i - 1] - c, axis=1)) >= 1.2:
                p[i] = c
                break
    return (p - p.mean(0)).astype(np.float32)


def main():
    if not orc.have_ref():
        raise SystemExit("oracle/_ref/libcdkref.so missing: run `make -C oracle ref` first")
    rs = np.random.default_rng(20260101)
    d = {}
    # A1: QuaternionUniformSampling, 3 successive calls of 40 on one mt19937_64.
    d["qus_seed"] = np.array([987654321], np.uint64)
    d["qus_out"] = orc.quaternion_uniform_sampling(987654321, 3, 40, lib="ref")
    # A3/A4: rotation and AABB of a 24-atom walk under 64 quaternions,
    # including identity, q and -q, axis-aligned and unnormalised ones.
    x = walk(rs, 24)
    q = unit_quats(rs, 58)
    extra = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0.5, 0.5, 0.5, 0.5], [0.9, 0.1, -0.2, 0.3],
                      [0, 0, 0, 0]], np.float32)
    q = np.concatenate([q, -q[:1], extra[:5]], axis=0)
    d["rot_coords"] = x
    d["rot_quats"] = q
    rot = orc.rotate(q, x, lib="ref")
    d["rot_out"] = rot
    mn, mx = orc.min_max_coors(rot, lib="ref")
    d["min_out"], d["max_out"] = mn, mx
    # signed zeros / ties for the min-max first-occurrence rule
    tie = np.array([[[0.0, 1.0, -0.0], [-0.0, 1.0, 0.0], [2.0, -1.0, -0.0]]], np.float32)
    tmn, tmx = orc.min_max_coors(tie, lib="ref")
    d["tie_coords"], d["tie_min"], d["tie_max"] = tie, tmn, tmx
    # A5: channel assignment with exact mid-point ties and off-set radii.
    ar = np.concatenate([rs.choice(PROBE, 30), rs.uniform(1.0, 2.5, 20).astype(np.float32),
                         np.array([1.35, 1.6, 1.75, 2.15, 1.0, 3.0], np.float32)]).astype(np.float32)
    nu, used, ari = orc.idx_of_atoms_for_vdw_radius(ar, PROBE, lib="ref")
    d["idx_radii"], d["idx_num_used"], d["idx_used"], d["idx_ari"] = ar, np.array([nu]), used, ari
    # A7: FillLigandGrid on 3 rotations of a 20-atom ligand, 28^3 grid, h = 0.7.
    xl = walk(rs, 20)
    ql = unit_quats(rs, 3)
    charge = (0.25 * rs.standard_normal(20)).astype(np.float32)
    charge -= charge.mean()
    eps = rs.uniform(0.05, 0.2, 20).astype(np.float32)
    lr = rs.choice(PROBE, 20).astype(np.float32)
    nu, used, ari = orc.idx_of_atoms_for_vdw_radius(lr, PROBE, lib="ref")
    rotl = orc.rotate(ql, xl, lib="ref")
    mnl, _ = orc.min_max_coors(rotl, lib="ref")
    dims = (28, 28, 28)
    g = orc.fill_ligand_grid(rotl, mnl, charge, eps, nu, used, ari, 8, dims, 0.7, lib="ref")
    nz = np.flatnonzero(g)
    d.update(fill_coords=xl, fill_quats=ql, fill_rot=rotl, fill_min=mnl, fill_charge=charge, fill_eps=eps,
             fill_radius=lr, fill_num_used=np.array([nu]), fill_used=used, fill_ari=ari, fill_dims=np.array(dims),
             fill_spacing=np.array([0.7], np.float32), fill_shape=np.array(g.shape), fill_nz_idx=nz.astype(np.int64),
             fill_nz_val=g.reshape(-1)[nz])
    # A2: FilterQuaternions with a grid length that rejects some rotations.
    xf = walk(rs, 30)
    qf = unit_quats(rs, 200)
    ext = np.ptp(orc.rotate(qf, xf, lib="ref"), axis=1)
    lens = np.percentile(ext, 85, axis=0).astype(np.float32)
    d.update(filt_coords=xf, filt_quats=qf, filt_lens=lens, filt_out=orc.filter_quaternions(xf, qf, lens, lib="ref"))
    np.savez_compressed(OUT, **d)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")
    grid_files()


def _num(v, style):
    """One value in one of several spellings a grid file may hold."""
    if style == 0:
        return "%.9g" % v
    if style == 1:
        return "%.6E" % v
    if style == 2:
        return "%+.7f" % v
    s = "%.5g" % v
    return s.replace("0.", ".", 1) if s.startswith("0.") else s.replace("-0.", "-.", 1)


def grid_files():
    """ReadGrids.cpp:21-66 layout: 3 title lines, the header (numGrids X Y Z
    mid len spacing restraint_k), numGrids-1 radii, then per grid one
    separator line and X*Y*Z lines "x y z value", z fastest."""
    rs = np.random.default_rng(20260202)
    d = {}
    cases = {
        # LF, one header line, one radii line
        "plain": dict(dims=(4, 5, 6), ng=3, eol="\n", sep=" ", trail="", split=False),
        # CRLF, trailing blanks and tabs, the header over two lines, tab separators
        "crlf": dict(dims=(3, 4, 5), ng=4, eol="\r\n", sep="\t ", trail="  \t", split=True),
    }
    for name, c in cases.items():
        X, Y, Z = c["dims"]
        ng = c["ng"]
        vals = rs.standard_normal((ng, X, Y, Z)) * np.exp(rs.uniform(-12, 6, (ng, X, Y, Z)))
        vals.reshape(-1)[:3] = [0.0, -0.0, 3.0]
        vals = vals.astype(np.float32)
        radii = rs.choice(PROBE, ng - 1, replace=False)
        eol, sp, tr = c["eol"], c["sep"], c["trail"]
        head = [ng, X, Y, Z, 1.25, -2.5, 3.75, (X - 1) * 0.5, (Y - 1) * 0.5, (Z - 1) * 0.5, 0.5, 10.0]
        hs = [str(v) for v in head]
        lines = ["* grid file " + name + tr, "* second title line", "*"]
        if c["split"]:
            lines += [sp.join(hs[:4]) + tr, sp.join(hs[4:]) + tr]
        else:
            lines += [sp.join(hs) + tr]
        lines += [sp.join("%.2f" % r for r in radii) + tr]
        k = 0
        for g in range(ng):
            lines.append("grid %d" % g + tr)
            for ix in range(X):
                for iy in range(Y):
                    for iz in range(Z):
                        v = float(vals[g, ix, iy, iz])
                        xyz = sp.join("%.3f" % t for t in (1.25 + ix * 0.5, -2.5 + iy * 0.5, 3.75 + iz * 0.5))
                        lines.append(xyz + sp + _num(v, k % 4) + tr)
                        k += 1
        path = os.path.join(HERE, "grids_%s.txt" % name)
        with open(path, "w", newline="") as f:
            f.write(eol.join(lines) + eol)
        rv, rr, rm, rl, rh, rk = orc.read_grids_ref(path)
        d[name + "_values"] = rv
        d[name + "_radii"] = rr
        d[name + "_mid"] = rm
        d[name + "_len"] = rl
        d[name + "_spacing"] = np.array([rh], np.float32)
        d[name + "_restraint_k"] = np.array([rk], np.float32)
    np.savez_compressed(GRID_OUT, **d)
    print("wrote", GRID_OUT, os.path.getsize(GRID_OUT), "bytes")


if __name__ == "__main__":
    main()
