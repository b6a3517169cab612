"""Seeded synthetic workloads of the five benchmark configurations.

Every array comes from the library's own generators (cdk_synth_*), so the
same seed gives the same inputs on every machine.  Conformers of a ligand each
draw their own quaternions from one std::mt19937_64 stream, in conformer
order, as TranRotaConfSearchNoMini.cu:321-324 redraws them per conformer.
"""
import numpy as np

from . import api

PROBE_RADII = np.array([1.2, 1.5, 1.7, 1.8, 1.9, 2.0, 2.1, 2.2], np.float32)

# name: (protein atoms, box A, grid edge, spacing A)
PROTEINS = {
    "p48": (2000, 24.0, 48, 0.5),
    "p64": (3000, 32.0, 64, 0.5),
    "p96": (8000, 48.0, 96, 0.5),
    "p128": (8000, 48.0, 128, 0.375),
}
PROTEIN_SEED = {"p48": 11, "p64": 12, "p96": 13, "p128": 14}


def protein_grids(name):
    natom, box, n, h = PROTEINS[name]
    vals, radii, mid, ln = api.synth_protein(PROTEIN_SEED[name], natom, box, n, h)
    return dict(values=vals, radii=radii, mid=mid, length=ln, spacing=h, n=n)


def ligand(seed, natom, nconf):
    x, q, e, r = api.synth_ligand(seed, natom)
    confs = api.synth_conformers(seed + 7919, x, nconf)
    return dict(charge=q, epsilon=e, radius=r, conformers=confs)


def chain_topology(natom):
    """Torsion topology of the synthetic ligands (a chain bonded i-(i+1)):
    rotor (b-1, b, b+1, b+2) for each inner bond b -> b+1 turns atoms b+2..
    (the bonds the synthetic conformers twist); 1-2, 1-3 and 1-4 pairs
    excluded from the clash test.  Inputs for generate_conformers."""
    bonds = np.arange(1, max(natom - 2, 1))
    rot = np.stack([bonds - 1, bonds, bonds + 1, bonds + 2], axis=1).astype(np.int32).reshape(-1, 4)
    mv = (np.arange(natom)[None, :] >= rot[:, 2:3] + 1).astype(np.uint8)
    sep = np.abs(np.arange(natom)[:, None] - np.arange(natom)[None, :])
    return rot, mv, (sep <= 3).astype(np.uint8)


def quaternion_sets(seed, nconf, nquat):
    """[nconf][nquat][4]: successive QuaternionUniformSampling draws on one mt19937_64."""
    rng = api.Rng(seed)
    out = np.stack([api.QuaternionUniformSampling(rng, nquat) for _ in range(nconf)])
    rng.close()
    return out


def config1():
    """48^3 x 9, one 24-atom ligand, 1 conformer x 500 quaternions (CPU-runnable case)."""
    lig = ligand(101, 24, 1)
    return dict(protein="p48", ligands=[dict(lig, quats=quaternion_sets(201, 1, 500))])


def config2(nconf=100, nquat=2000):
    """64^3 x 9, one 40-atom ligand, 100 conformers x 2000 quaternions."""
    lig = ligand(102, 40, nconf)
    return dict(protein="p64", ligands=[dict(lig, quats=quaternion_sets(202, nconf, nquat))])


def config3(nlig=64, nquat=1000, conf_range=(10, 100), atom_range=(15, 60)):
    """64 ligands, atoms U[15,60], conformers U[10,100], 1000 quaternions each."""
    rs = np.random.default_rng(303)
    ligs = []
    for i in range(nlig):
        na = int(rs.integers(atom_range[0], atom_range[1] + 1))
        nc = int(rs.integers(conf_range[0], conf_range[1] + 1))
        lig = ligand(1000 + i, na, nc)
        ligs.append(dict(lig, quats=quaternion_sets(2000 + i, nc, nquat)))
    return dict(protein="p64", ligands=ligs)


def serving_request(rs, seed):
    """One request of the serving stream: atoms U[15,80], conformers
    U[1,50], quaternions in {500, 2000, 5000}, protein 64^3 or 96^3."""
    na = int(rs.integers(15, 81))
    nc = int(rs.integers(1, 51))
    nq = int(rs.choice([500, 2000, 5000]))
    prot = int(rs.integers(0, 2))
    lig = ligand(seed, na, nc)
    return dict(lig, quats=quaternion_sets(seed + 1, nc, nq), protein_index=prot)


def config5(nconf=50, nquat=3000):
    """128^3 x 9 at 0.375 A, one 50-atom ligand, 50 conformers x 3000 quaternions."""
    lig = ligand(105, 50, nconf)
    return dict(protein="p128", ligands=[dict(lig, quats=quaternion_sets(205, nconf, nquat))])
