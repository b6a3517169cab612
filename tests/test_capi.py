"""C-ABI and host logic without a GPU: the library loads, exports every
symbol include/cdk.h declares, reports errors as status codes, and its
host-side producers (quaternion sampling, channel assignment, synthetic
inputs, grid file I/O) match the oracle / the reference's own outputs."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import cdocker_amd as cdk
from cdocker_amd import _lib, api
from oracle import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = np.load(os.path.join(ROOT, "tests", "golden", "ref_vectors.npz"))
PROBE = np.array([1.2, 1.5, 1.7, 1.8, 1.9, 2.0, 2.1, 2.2], np.float32)


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "cdk.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(cdk_[a-z0-9_]+)\(", text, re.M)))


def test_every_declared_symbol_is_exported_and_bound():
    names = declared_symbols()
    assert len(names) >= 40
    for n in names:
        assert hasattr(_lib.lib, n), n
    assert names == _lib.DECLARED


def test_status_codes_without_device():
    h = C.c_void_p()
    rc = _lib.lib.cdk_context_create(12345, C.byref(h))
    assert rc in (_lib.ERR_INVALID, _lib.ERR_HIP)
    assert len(_lib.last_error()) > 0
    assert _lib.lib.cdk_context_create(0, None) == _lib.ERR_INVALID


def test_quaternion_sampling_matches_reference_stream():
    rng = cdk.Rng(int(GOLD["qus_seed"][0]))
    got = np.stack([cdk.QuaternionUniformSampling(rng, 40) for _ in range(3)])
    assert np.array_equal(got.view(np.uint32), GOLD["qus_out"].view(np.uint32))
    rng.close()
    norms = np.linalg.norm(got.reshape(-1, 4).astype(np.float64), axis=1)
    assert np.allclose(norms, 1, atol=1e-6)


def test_idx_of_atoms_matches_reference():
    nu, used, ari = cdk.GetIdxOfAtomsForVdwRadius(GOLD["idx_radii"], PROBE)
    assert nu == int(GOLD["idx_num_used"][0])
    assert np.array_equal(used, GOLD["idx_used"]) and np.array_equal(ari, GOLD["idx_ari"])


def test_sort_index_ties_by_position():
    v = np.array([3.0, 1.0, 2.0, 1.0, np.inf, -1.0], np.float32)
    assert cdk.sort_index(v).tolist() == [5, 1, 3, 2, 0, 4]


def test_synthetic_generators_are_deterministic_and_valid():
    x1, q1, e1, r1 = cdk.synth_ligand(5, 40)
    x2, q2, e2, r2 = cdk.synth_ligand(5, 40)
    assert np.array_equal(x1, x2) and np.array_equal(q1, q2)
    d = np.linalg.norm(x1[1:] - x1[:-1], axis=1)
    assert np.allclose(d, 1.5, atol=1e-4)
    dm = np.linalg.norm(x1[:, None] - x1[None], axis=2)
    assert dm[np.triu_indices(40, 2)].min() >= 1.2 - 1e-4
    assert abs(float(q1.sum())) < 1e-4
    assert set(np.unique(r1)).issubset(set(PROBE.tolist()))
    confs = cdk.synth_conformers(6, x1, 5)
    assert np.array_equal(confs[0] - confs[0].mean(0), confs[0])  or np.allclose(confs[0].mean(0), 0, atol=1e-5)
    for c in confs:  # torsions keep bond lengths
        assert np.allclose(np.linalg.norm(c[1:] - c[:-1], axis=1), 1.5, atol=1e-3)
    v, radii, mid, ln = cdk.synth_protein(3, 300, 16.0, 32, 0.5)
    assert v.shape == (9, 32, 32, 32) and np.all(v[:8] <= 3.0) and np.all(np.abs(v[8]) <= 20.0)
    assert np.array_equal(radii, PROBE) and np.allclose(ln, 31 * 0.5)


def test_chain_topology_turns_the_synthetic_torsions():
    """workloads.chain_topology, the conformer-generation inputs the bench
    uses for the synthetic chain ligands: each rotor's four atoms are
    consecutive chain atoms, a rotor about bond b -> b+1 moves exactly atoms
    b+2.., turning it keeps every bond length and the far side rigid, and the
    exclusions are the 1-2, 1-3 and 1-4 pairs (the same as the tests'
    oracle helper)."""
    from cdocker_amd import workloads
    for n in (4, 5, 12, 40):
        rot, mv, ex = workloads.chain_topology(n)
        r2, m2, e2 = orc.chain_rotors(n)
        assert np.array_equal(rot, r2) and np.array_equal(mv, m2) and np.array_equal(ex, e2)
        assert rot.shape == (n - 3, 4) and (np.diff(rot, axis=1) == 1).all()
    x, _, _, _ = cdk.synth_ligand(5, 12)
    rot, mv, _ = workloads.chain_topology(12)
    x = x.astype(np.float64)
    for (a, b, c, d), m in zip(rot, mv):
        axis = (x[c] - x[b]) / np.linalg.norm(x[c] - x[b])
        th = 0.7
        k = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
        R = np.eye(3) + np.sin(th) * k + (1 - np.cos(th)) * (k @ k)
        y = x.copy()
        y[m.astype(bool)] = (x[m.astype(bool)] - x[c]) @ R.T + x[c]
        assert np.allclose(np.linalg.norm(y[1:] - y[:-1], axis=1), np.linalg.norm(x[1:] - x[:-1], axis=1))
        moved = np.flatnonzero(np.linalg.norm(y - x, axis=1) > 1e-9)
        assert moved.min() >= c + 1  # atoms up to c stay put
        far = m.astype(bool)
        dfar = np.linalg.norm(y[far][:, None] - y[far][None], axis=2)
        assert np.allclose(dfar, np.linalg.norm(x[far][:, None] - x[far][None], axis=2))


def test_grid_file_round_trip(tmp_path):
    rs = np.random.default_rng(0)
    vals = rs.standard_normal((3, 4, 5, 6)).astype(np.float32)
    p = tmp_path / "g.txt"
    cdk.write_grids(p, vals, np.array([1.5, 2.0], np.float32), np.array([1, 2, 3], np.float32),
                    np.array([1.5, 2, 2.5], np.float32), 0.5, 7.0)
    v2, r2, mid, ln, sp, rk = cdk.read_grids(p)
    assert np.array_equal(v2, vals) and np.allclose(r2, [1.5, 2.0]) and sp == 0.5 and rk == 7.0
    bad = tmp_path / "bad.txt"
    bad.write_text("a\nb\nc\n2 4 4\n")
    with pytest.raises(cdk.CdkError):
        cdk.read_grids(bad)


@pytest.mark.parametrize("name", ["plain", "crlf"])
def test_grid_reader_matches_reference_reader(name):
    """cdk_read_grids on grid files with LF and CRLF endings, trailing blanks
    and tabs, a header over two lines and assorted number spellings gives
    what the reference's own ReadGrids.cpp:6-68 read from the same files,
    bit for bit (signed zeros included)."""
    gold = np.load(os.path.join(ROOT, "tests", "golden", "grids_ref.npz"))
    v, radii, mid, ln, sp, rk = cdk.read_grids(os.path.join(ROOT, "tests", "golden", "grids_%s.txt" % name))
    assert np.array_equal(v.view(np.uint32), gold[name + "_values"].view(np.uint32))
    assert np.array_equal(radii.view(np.uint32), gold[name + "_radii"].view(np.uint32))
    assert np.array_equal(mid, gold[name + "_mid"]) and np.array_equal(ln, gold[name + "_len"])
    assert np.float32(sp) == gold[name + "_spacing"][0] and np.float32(rk) == gold[name + "_restraint_k"][0]
    assert np.signbit(v.reshape(-1)[1]) and v.reshape(-1)[2] == 3.0


def test_write_poses_files_and_energy_rows(tmp_path):
    """conformer_<c>_<i>.pdb files and energy.txt rows as the drivers write
    them (TranRotaConfSearchNoMini.cu:489-502); runs on the host only."""
    rs = np.random.default_rng(2)
    poses = rs.normal(0, 5, size=(3, 4, 3))
    energy = np.array([-12.5, -11.25, 3.0e-5], np.float32)
    names = ["C1", "N2", "O3", "CL4"]
    elements = ["C", "N", "O", "Cl"]
    epath = tmp_path / "energy.txt"
    epath.write_text("file,conformer,rank,energy\n")
    cdk.write_poses(tmp_path, 7, poses, energy, names, elements, str(epath))
    rows = epath.read_text().splitlines()
    assert rows == ["file,conformer,rank,energy", "conformer_7_0.pdb,7,0,-12.5", "conformer_7_1.pdb,7,1,-11.25",
                    "conformer_7_2.pdb,7,2,3e-05"]
    for i in range(3):
        lines = (tmp_path / ("conformer_7_%d.pdb" % i)).read_text().splitlines()
        atoms = [l for l in lines if l.startswith("HETATM")]
        assert len(atoms) == 4 and lines[-1] == "END"
        for a, l in enumerate(atoms):
            assert len(l) == 78
            assert int(l[6:11]) == a + 1 and l[12:16].strip() == names[a] and l[17:20] == "LIG"
            xyz = [float(l[30:38]), float(l[38:46]), float(l[46:54])]
            assert np.allclose(xyz, poses[i, a], atol=5e-4)
            assert l[76:78].strip() == elements[a]


def test_synthetic_protein_matches_grid_builder_oracle():
    """The host synthetic recipe (cell-list order) equals the oracle's
    brute-force grid builder on the same atoms."""
    n, h = 32, 0.5
    x, q, e, r = cdk.synth_protein_atoms(5, 300, 16.0)
    v, radii, mid, ln = cdk.synth_protein(5, 300, 16.0, n, h)
    gm = np.float64(np.float32(mid[0]) - np.float32(ln[0]) / np.float32(2))
    g = orc.build_grids(x, q, e, r, radii, n, np.full(3, gm), h)
    assert np.array_equal(g, v)
