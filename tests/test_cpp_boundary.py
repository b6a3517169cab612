"""The C-ABI from a C++ caller: tests/cpp/nomini_block.cpp, the
TranRotaConfSearchNoMini.cu:156-487 block written against include/cdk.h and
the reference-signature adapter include/cdk_reference.hpp (g++ -std=c++11,
linked against libcdk.so; built by tests/cpp/Makefile from build()).

On the host: Rotate, GetMinCoors, GetMaxCoors, GetIdxOfAtomsForVdwRadius and
QuaternionUniformSampling called with the reference's signatures give the
reference's own outputs (golden vectors of the compiled reference sources)
bit for bit, and the C-ABI reports bad arguments and a missing device as
status codes.  On the GPU: the driver loop's results equal cdk_search /
cdk_rank_poses through the Python binding."""
import os
import struct
import subprocess

import numpy as np
import pytest

import cdocker_amd as cdk

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "_build", "nomini_block")
GOLD = np.load(os.path.join(ROOT, "tests", "golden", "ref_vectors.npz"))
PROBE = np.array([1.2, 1.5, 1.7, 1.8, 1.9, 2.0, 2.1, 2.2], np.float32)


def exe():
    if not os.path.exists(EXE):
        subprocess.run(["make", "-C", os.path.join(ROOT, "tests", "cpp")], check=True)
    return EXE


def put(f, a):
    a = np.ascontiguousarray(a)
    f.write(struct.pack("<i", a.size))
    f.write(a.tobytes())


def records(path, dtypes):
    out = []
    with open(path, "rb") as f:
        for dt in dtypes:
            (n,) = struct.unpack("<i", f.read(4))
            out.append(np.frombuffer(f.read(n * np.dtype(dt).itemsize), dt).copy())
    return out


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def test_reference_signatures_on_host_match_golden(tmp_path):
    qus = GOLD["qus_out"]
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        put(f, GOLD["rot_quats"].astype(np.float32))
        put(f, GOLD["rot_coords"].astype(np.float32))
        put(f, GOLD["idx_radii"].astype(np.float32))
        put(f, PROBE)
        put(f, np.array([int(GOLD["qus_seed"][0]), qus.shape[0], qus.shape[1]], np.int64))
    out = tmp_path / "out.bin"
    r = subprocess.run([exe(), "host", str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    rot, mn, mx, nu, used, ari, q, codes = records(
        out, [np.float32, np.float32, np.float32, np.int32, np.int32, np.int32, np.float32, np.int32])
    nq, na = GOLD["rot_quats"].shape[0], GOLD["rot_coords"].shape[0]
    assert np.array_equal(bits(rot).reshape(nq, na, 3), bits(GOLD["rot_out"]))
    assert np.array_equal(bits(mn).reshape(nq, 3), bits(GOLD["min_out"]))
    assert np.array_equal(bits(mx).reshape(nq, 3), bits(GOLD["max_out"]))
    assert nu[0] == int(GOLD["idx_num_used"][0]) and np.array_equal(used, GOLD["idx_used"])
    assert np.array_equal(ari, GOLD["idx_ari"])
    assert np.array_equal(bits(q).reshape(qus.shape), bits(qus))
    # status codes: bad counts, missing device, bad arguments; a message set
    assert codes[0] == 1 and codes[1] == 1 and codes[2] != 0 and codes[3] == 0 and codes[4] == 1 and codes[5] == 1


@pytest.mark.gpu
def test_driver_loop_on_gpu_matches_python_binding(tmp_path):
    out = tmp_path / "search.bin"
    r = subprocess.run([exe(), "search", str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    recs = records(out, [np.float32, np.float32, np.int32, np.float32, np.int32, np.float64, np.float64] * 3)
    ctx = cdk.Context(0)
    vals, radii, mid, ln = cdk.synth_protein(12, 3000, 32.0, 64, 0.5)
    prot = cdk.Protein(ctx, vals, radii, 0.5, mid, ln)
    x, q, e, rad = cdk.synth_ligand(102, 40)
    confs = cdk.synth_conformers(102 + 7919, x, 3)
    lig = cdk.Ligand(prot, q, e, rad)
    try:
        for c in range(3):
            qu, E, idx, mc, rank, pose, mass = recs[7 * c: 7 * c + 7]
            qu = qu.reshape(-1, 4)
            assert qu.shape[0] > 50
            E1, i1, m1 = cdk.search(ctx, prot, lig, confs[c], qu)
            assert np.array_equal(bits(E1[0]), bits(E)) and np.array_equal(i1[0].reshape(-1), idx)
            assert np.array_equal(bits(m1[0].reshape(-1)), bits(mc))
            rot = ctx.Rotate(qu, confs[c])
            rk, ps = cdk.rank_poses(prot, rot, E1[0], i1[0], m1[0], 5)
            assert np.array_equal(rk, rank) and np.array_equal(ps.reshape(-1), pose)
            # the best pose's spread mass: sum of sqrt(eps) over atoms plus the (zero) total charge
            assert abs(mass[0] - float(np.sqrt(e.astype(np.float64)).sum() + q.astype(np.float64).sum())) < 1e-3
    finally:
        lig.close()
        prot.close()
        ctx.close()
