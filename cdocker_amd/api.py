"""Python mirror of the reference's pose-search call surface over libcdk.so.

Names follow the reference's functions (Rotate, GetMinCoors, FillLigandGrid,
QuaternionUniformSampling, FilterQuaternions, GetIdxOfAtomsForVdwRadius,
sort_index, TranRotaConfSearch) with numpy arrays in place of raw pointers.
Every compute call runs on the GPU through the C-ABI.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import lib, check, ENGINE_REFERENCE, ENGINE_FUSED, CdkError  # noqa: F401


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        a = a.reshape(shape)
    return a


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _fp(a):
    return a.ctypes.data_as(_lib.FP) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(_lib.IP) if a is not None else None


def _dp(a):
    return a.ctypes.data_as(_lib.DP) if a is not None else None


class Context:
    """Device context (one HIP stream, work arrays, twiddle tables)."""

    def __init__(self, device=0, **options):
        h = C.c_void_p()
        check(lib.cdk_context_create(int(device), C.byref(h)), "cdk_context_create")
        self.h = h
        for k, v in options.items():
            self.set_option(k, v)

    def set_option(self, key, value):
        check(lib.cdk_context_set_option(self.h, key.encode(), int(value)), "set_option " + key)

    def get_option(self, key):
        v = C.c_int64()
        check(lib.cdk_context_get_option(self.h, key.encode(), C.byref(v)), "get_option " + key)
        return v.value

    def stage_times(self, reset=False):
        ms = np.zeros(8, np.float64)
        n = np.zeros(8, np.int32)
        check(lib.cdk_context_stage_times(self.h, _dp(ms), _ip(n), int(bool(reset))))
        return ms, n

    def close(self):
        if self.h:
            check(lib.cdk_context_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- reference pose-preparation surface (device) -------------------
    def Rotate(self, quats, coords):
        """Rotate.cpp:4-27 for every (quaternion, atom) -> [nq][natom][3]."""
        q = _f32(quats, (-1, 4))
        x = _f32(coords, (-1, 3))
        out = np.empty((q.shape[0], x.shape[0], 3), np.float32)
        check(lib.cdk_rotate(self.h, q.shape[0], x.shape[0], _fp(q), _fp(x), _fp(out)), "cdk_rotate")
        return out

    def rotate_aabb(self, quats, coords, grid_edge, spacing):
        """Rotation, GetMinCoors/GetMaxCoors extents and the translation window."""
        q = _f32(quats, (-1, 4))
        x = _f32(coords, (-1, 3))
        nq, na = q.shape[0], x.shape[0]
        rot = np.empty((nq, na, 3), np.float32)
        mn = np.empty((nq, 3), np.float32)
        ln = np.empty((nq, 3), np.float32)
        win = np.empty((nq, 3), np.int32)
        check(lib.cdk_rotate_aabb(self.h, nq, na, _fp(q), _fp(x), int(grid_edge), float(spacing), _fp(rot),
                                  _fp(mn), _fp(ln), _ip(win)), "cdk_rotate_aabb")
        return rot, mn, ln, win

    def FillLigandGrid(self, coors, mincoors, charge, epsilon, num_vdw_used, idx_vdw_used, atoms_of_radius, xdim,
                       ydim, zdim, spacing, grid=None):
        """FillLigandGrid.cpp:7-120; adds into `grid` ([nq][C][X][Y][Z], zeros if None)."""
        co = _f32(coors, (-1,))
        mc = _f32(mincoors, (-1, 3))
        nq = mc.shape[0]
        q = _f32(charge)
        e = _f32(epsilon)
        natom = q.shape[0]
        used = _i32(idx_vdw_used)
        nv = len(atoms_of_radius)
        offs = np.zeros(nv + 1, np.int32)
        flat = []
        for r, lst in enumerate(atoms_of_radius):
            offs[r] = len(flat)
            flat.extend(int(a) for a in lst)
        offs[nv] = len(flat)
        flat = _i32(flat if flat else [0])
        C_ = int(num_vdw_used) + 1
        if grid is None:
            grid = np.zeros((nq, C_, xdim, ydim, zdim), np.float32)
        g = np.ascontiguousarray(grid, dtype=np.float32)
        check(lib.cdk_fill_ligand_grid(self.h, nq, natom, _fp(co), _fp(mc), _fp(q), _fp(e), int(num_vdw_used),
                                       _ip(used), nv, _ip(offs), _ip(flat), xdim, ydim, zdim, float(spacing),
                                       _fp(g)), "cdk_fill_ligand_grid")
        return g

    def fft_r2c_3d(self, grids):
        g = _f32(grids)
        n = g.shape[-1]
        b = g.size // (n * n * n)
        out = np.empty((b, n, n, n // 2 + 1), np.complex64)
        check(lib.cdk_fft_r2c_3d(self.h, n, b, _fp(g), out.ctypes.data_as(_lib.FP)), "cdk_fft_r2c_3d")
        return out

    def fft_c2r_3d(self, spectra):
        s = np.ascontiguousarray(spectra, dtype=np.complex64)
        n = s.shape[-2]
        b = s.size // (n * n * (n // 2 + 1))
        out = np.empty((b, n, n, n), np.float32)
        check(lib.cdk_fft_c2r_3d(self.h, n, b, s.ctypes.data_as(_lib.FP), _fp(out)), "cdk_fft_c2r_3d")
        return out

    def ConjMult(self, prot, lig):
        """kernel.cu:5-14: lig[q][c] <- conj(lig[q][c]) * prot[c]."""
        p = np.ascontiguousarray(prot, dtype=np.complex64)
        l = np.array(lig, dtype=np.complex64, copy=True, order="C")
        nch = p.shape[0]
        H = p.size // nch
        nq = l.size // (nch * H)
        check(lib.cdk_conj_mult(self.h, nq, nch, H, p.ctypes.data_as(_lib.FP), l.ctypes.data_as(_lib.FP)),
              "cdk_conj_mult")
        return l

    def SumGrids(self, lig, nch, G):
        """kernel.cu:16-30: sum over channels divided by sqrtf(G)."""
        l = np.ascontiguousarray(lig, dtype=np.complex64)
        nq = l.shape[0]
        H = l.size // (nq * nch)
        out = np.empty((nq, H), np.complex64)
        check(lib.cdk_sum_grids(self.h, nq, nch, H, int(G), l.ctypes.data_as(_lib.FP), out.ctypes.data_as(_lib.FP)),
              "cdk_sum_grids")
        return out

    def argmin_window(self, grids, windows):
        g = _f32(grids)
        n = g.shape[-1]
        nq = g.size // (n * n * n)
        w = _i32(windows).reshape(nq, 3)
        e = np.empty(nq, np.float32)
        idx = np.empty((nq, 3), np.int32)
        check(lib.cdk_argmin_window(self.h, nq, n, _fp(g), _ip(w), _fp(e), _ip(idx)), "cdk_argmin_window")
        return e, idx

    def FilterQuaternions(self, coor, quats, len_xyz, max_used, shuffle_seed=0):
        """FilterQuaternions.cpp:11-83 (extent test on the device)."""
        x = _f32(coor, (-1, 3))
        q = _f32(quats, (-1, 4))
        out = np.empty_like(q)
        n = C.c_int()
        check(lib.cdk_filter_quaternions(self.h, _fp(x), x.shape[0], q.shape[0], _fp(q), float(len_xyz[0]),
                                         float(len_xyz[1]), float(len_xyz[2]), int(max_used), int(shuffle_seed),
                                         C.byref(n), _fp(out)), "cdk_filter_quaternions")
        return out[: n.value].copy()


class Protein:
    """Resident potential grids and their spectra (NoMini:177-199)."""

    def __init__(self, ctx, values, radii, spacing, mid=None, length=None):
        v = _f32(values)
        ng, n = v.shape[0], v.shape[-1]
        r = _f32(radii)
        mid = _f32(mid if mid is not None else [0, 0, 0])
        length = _f32(length if length is not None else [(n - 1) * spacing] * 3)
        h = C.c_void_p()
        check(lib.cdk_protein_create(ctx.h, ng, n, n, n, _fp(mid), _fp(length), float(spacing), _fp(r), _fp(v),
                                     C.byref(h)), "cdk_protein_create")
        self.h, self.ctx = h, ctx
        self.num_grids, self.n, self.spacing = ng, n, float(spacing)
        self.mid, self.length, self.radii = mid, length, r

    @classmethod
    def build(cls, ctx, xyz, charge, epsilon, radius, probe_radii, n, spacing, mid, length, lj_cutoff=8.0,
              elec_cutoff=12.0, lj_cap=3.0, elec_cap=20.0):
        """Potential grids built on the device from atoms (cdk_protein_build)."""
        x = np.ascontiguousarray(np.asarray(xyz, np.float64).reshape(-1, 3))
        q = np.ascontiguousarray(np.asarray(charge, np.float64))
        e = np.ascontiguousarray(np.asarray(epsilon, np.float64))
        r = np.ascontiguousarray(np.asarray(radius, np.float64))
        pr = _f32(probe_radii)
        mid = _f32(mid)
        length = _f32(length)
        h = C.c_void_p()
        check(lib.cdk_protein_build(ctx.h, x.shape[0], _dp(x), _dp(q), _dp(e), _dp(r), pr.shape[0], _fp(pr), int(n),
                                    float(spacing), _fp(mid), _fp(length), float(lj_cutoff), float(elec_cutoff),
                                    float(lj_cap), float(elec_cap), C.byref(h)), "cdk_protein_build")
        self = cls.__new__(cls)
        self.h, self.ctx = h, ctx
        self.num_grids, self.n, self.spacing = pr.shape[0] + 1, int(n), float(spacing)
        self.mid, self.length, self.radii = mid, length, pr
        return self

    def grids(self):
        n = self.n
        out = np.empty((self.num_grids, n, n, n), np.float32)
        check(lib.cdk_protein_grids(self.h, _fp(out)), "cdk_protein_grids")
        return out

    def spectrum(self, channel):
        n = self.n
        out = np.empty((n, n, n // 2 + 1), np.complex64)
        check(lib.cdk_protein_spectrum(self.h, int(channel), out.ctypes.data_as(_lib.FP)))
        return out

    def device_bytes(self):
        b = C.c_int64()
        check(lib.cdk_protein_bytes(self.h, C.byref(b)))
        return b.value

    def grid_min(self):
        """gridMin = mid - len / 2 in float (NoMini:149-151)."""
        return (self.mid - self.length / np.float32(2)).astype(np.float64)

    def close(self):
        if self.h:
            check(lib.cdk_protein_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Ligand:
    def __init__(self, protein, charge, epsilon, radius):
        self.charge = _f32(charge)
        self.epsilon = _f32(epsilon)
        self.radius = _f32(radius)
        self.natom = self.charge.shape[0]
        h = C.c_void_p()
        check(lib.cdk_ligand_create(protein.h, self.natom, _fp(self.charge), _fp(self.epsilon), _fp(self.radius),
                                    C.byref(h)), "cdk_ligand_create")
        self.h, self.protein = h, protein

    def channels(self):
        nu = C.c_int()
        used = np.zeros(self.protein.num_grids - 1, np.int32)
        ari = np.zeros(self.natom, np.int32)
        check(lib.cdk_ligand_channels(self.h, C.byref(nu), _ip(used), _ip(ari)))
        return nu.value, used[: nu.value].copy(), ari

    def close(self):
        if self.h:
            check(lib.cdk_ligand_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def search(ctx, protein, ligand, conformers, quats, engine=ENGINE_FUSED):
    """The TranRotaConfSearch hot loop (NoMini:297-464).

    conformers [nconf][natom][3]; quats [nconf][nquat][4] or [nquat][4]
    (shared).  Returns energy [nconf][nquat], index [nconf][nquat][3],
    mincoor [nconf][nquat][3].
    """
    x = _f32(conformers, (-1, ligand.natom, 3))
    q = _f32(quats)
    shared = q.ndim == 2
    nconf = x.shape[0]
    nquat = q.shape[-2]
    e = np.empty((nconf, nquat), np.float32)
    idx = np.empty((nconf, nquat, 3), np.int32)
    mc = np.empty((nconf, nquat, 3), np.float32)
    check(lib.cdk_search(ctx.h, protein.h, ligand.h, nconf, _fp(x), nquat, _fp(q), int(shared), int(engine), _fp(e),
                         _ip(idx), _fp(mc)), "cdk_search")
    return e, idx, mc


def rank_poses(protein, rotated, energy, index, mincoor, n_lowest):
    """sort_index + pose rebuild of one conformer (NoMini:466-487)."""
    rot = _f32(rotated)
    nq, na = rot.shape[0], rot.shape[1]
    e = _f32(energy)
    idx = _i32(index)
    mc = _f32(mincoor)
    n = min(int(n_lowest), nq)
    rank = np.empty(max(n, 1), np.int32)
    pose = np.empty((max(n, 1), na, 3), np.float64)
    check(lib.cdk_rank_poses(protein.h, na, nq, _fp(rot), _fp(e), _ip(idx), _fp(mc), int(n_lowest), _ip(rank),
                             _dp(pose)), "cdk_rank_poses")
    return rank[:n], pose[:n]


def rescore_poses(protein, ligand, poses, want_grad=False):
    """FP64 grid score of explicit poses [npose][natom][3] (A).  Returns
    (energy [npose], grad [npose][natom][3] or None, atoms on the grid)."""
    x = np.ascontiguousarray(np.asarray(poses, np.float64).reshape(-1, ligand.natom, 3))
    n = x.shape[0]
    e = np.empty(n, np.float64)
    g = np.empty_like(x) if want_grad else None
    ins = np.empty(n, np.int32)
    check(lib.cdk_rescore_poses(protein.h, ligand.h, n, _dp(x), _dp(e), _dp(g) if want_grad else None, _ip(ins)),
          "cdk_rescore_poses")
    return e, g, ins


def refine_poses(protein, ligand, poses, steps=50, off_grid_penalty=10.0):
    """Rigid-body descent on the grid score; returns (refined poses, energy)."""
    x = np.array(np.asarray(poses, np.float64).reshape(-1, ligand.natom, 3), order="C", copy=True)
    n = x.shape[0]
    e = np.empty(n, np.float64)
    check(lib.cdk_refine_poses(protein.h, ligand.h, n, _dp(x), int(steps), float(off_grid_penalty), _dp(e)),
          "cdk_refine_poses")
    return x, e


def write_poses(directory, conformer, poses, energy, atom_names=None, elements=None, energy_path=None):
    """conformer_<c>_<i>.pdb files + energy.txt rows (NoMini:489-502)."""
    x = np.ascontiguousarray(np.asarray(poses, np.float64))
    n, na = x.shape[0], x.shape[1]
    e = _f32(energy)

    def strs(v):
        if v is None:
            return None
        arr = (C.c_char_p * na)(*[str(t).encode() for t in v])
        return C.cast(arr, C.POINTER(C.c_char_p))

    check(lib.cdk_write_poses(str(directory).encode(), int(conformer), n, na, _dp(x), _fp(e), strs(atom_names),
                              strs(elements), energy_path.encode() if energy_path else None), "cdk_write_poses")


def grid_for_openmm(protein, channel):
    """(values [Z][Y][X] float64 -- x fastest -- , bounds_nm [6])."""
    n = protein.n
    v = np.empty((n, n, n), np.float64)
    b = np.empty(6, np.float64)
    check(lib.cdk_grid_for_openmm(protein.h, int(channel), _dp(v), _dp(b)), "cdk_grid_for_openmm")
    return v, b


def generate_conformers(ctx, coords, rotors, moving, exclude, ntrial, seed, clash_dist=1.2, max_keep=None):
    """Random-torsion conformers on the device; returns (conformers
    [n][natom][3] float64, trial index of each)."""
    x = np.ascontiguousarray(np.asarray(coords, np.float64).reshape(-1, 3))
    na = x.shape[0]
    r = np.ascontiguousarray(np.asarray(rotors, np.int32).reshape(-1, 4))
    mv = np.ascontiguousarray(np.asarray(moving, np.uint8).reshape(r.shape[0], na))
    ex = np.ascontiguousarray(np.asarray(exclude, np.uint8).reshape(na, na))
    keep = int(ntrial if max_keep is None else max_keep)
    out = np.empty((max(keep, 1), na, 3), np.float64)
    ti = np.empty(max(keep, 1), np.int32)
    n = C.c_int()
    check(lib.cdk_generate_conformers(ctx.h, na, _dp(x), r.shape[0], _ip(r), mv.ctypes.data_as(C.POINTER(C.c_uint8)),
                                      ex.ctypes.data_as(C.POINTER(C.c_uint8)), int(ntrial), C.c_uint64(seed),
                                      float(clash_dist), keep, _dp(out), _ip(ti), C.byref(n)),
          "cdk_generate_conformers")
    return out[:n.value], ti[:n.value]


class Batch:
    """Many ligands searched together in shared launches."""

    def __init__(self, ctx, protein):
        h = C.c_void_p()
        check(lib.cdk_batch_create(ctx.h, protein.h, C.byref(h)))
        self.h, self.jobs = h, []

    def add(self, ligand, conformers, quats):
        x = _f32(conformers, (-1, ligand.natom, 3))
        q = _f32(quats)
        shared = q.ndim == 2
        jid = C.c_int()
        check(lib.cdk_batch_add(self.h, ligand.h, x.shape[0], _fp(x), q.shape[-2], _fp(q), int(shared),
                                C.byref(jid)), "cdk_batch_add")
        self.jobs.append((x.shape[0], q.shape[-2]))
        return jid.value

    def run(self, engine=ENGINE_FUSED):
        check(lib.cdk_batch_run(self.h, int(engine)), "cdk_batch_run")

    def total_poses(self):
        n = C.c_int64()
        check(lib.cdk_batch_poses(self.h, C.byref(n)))
        return n.value

    def results(self, job):
        nconf, nquat = self.jobs[job]
        e = np.empty((nconf, nquat), np.float32)
        idx = np.empty((nconf, nquat, 3), np.int32)
        mc = np.empty((nconf, nquat, 3), np.float32)
        check(lib.cdk_batch_results(self.h, int(job), _fp(e), _ip(idx), _fp(mc)), "cdk_batch_results")
        return e, idx, mc

    def geometry(self, job):
        nconf, nquat = self.jobs[job]
        box = np.empty((nconf, nquat, 3), np.int32)
        win = np.empty((nconf, nquat, 3), np.int32)
        check(lib.cdk_batch_geometry(self.h, int(job), _ip(box), _ip(win)), "cdk_batch_geometry")
        return box, win

    def close(self):
        if self.h:
            check(lib.cdk_batch_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching search server over resident proteins."""

    def __init__(self, ctx, proteins, hbm_budget_bytes, max_batch_poses):
        arr = (C.c_void_p * len(proteins))(*[p.h.value for p in proteins])
        h = C.c_void_p()
        check(lib.cdk_server_create(ctx.h, len(proteins), C.cast(arr, _lib.PP), int(hbm_budget_bytes),
                                    int(max_batch_poses), C.byref(h)), "cdk_server_create")
        self.h, self.proteins, self.shapes = h, proteins, {}

    def submit(self, protein_index, charge, epsilon, radius, conformers, quats):
        q_ = _f32(charge)
        e_ = _f32(epsilon)
        r_ = _f32(radius)
        na = q_.shape[0]
        x = _f32(conformers, (-1, na, 3))
        q = _f32(quats)
        shared = q.ndim == 2
        rid = C.c_int64()
        check(lib.cdk_server_submit(self.h, int(protein_index), na, _fp(q_), _fp(e_), _fp(r_), x.shape[0], _fp(x),
                                    q.shape[-2], _fp(q), int(shared), C.byref(rid)), "cdk_server_submit")
        self.shapes[rid.value] = (x.shape[0], q.shape[-2])
        return rid.value

    def result(self, rid, wait_ms=-1):
        """Returns (energy, index, mincoor, latency_s) or None while pending."""
        nconf, nquat = self.shapes.get(rid, (0, 0))  # unknown or collected ids: the library reports it
        e = np.empty((nconf, nquat), np.float32)
        idx = np.empty((nconf, nquat, 3), np.int32)
        mc = np.empty((nconf, nquat, 3), np.float32)
        lat = C.c_double()
        rc = lib.cdk_server_result(self.h, C.c_int64(rid), int(wait_ms), _fp(e), _ip(idx), _fp(mc), C.byref(lat))
        if rc == _lib.ERR_PENDING:
            return None
        self.shapes.pop(rid, None)  # the server released the request
        check(rc, "cdk_server_result")
        return e, idx, mc, lat.value

    def stats(self):
        keys = ["requests_done", "poses_done", "launch_sets", "mean_batch_poses", "busy_s", "admitted_bytes",
                "queued", "peak_admitted_bytes", "held_requests", "sets_in_flight"]
        s = np.zeros(len(keys), np.float64)
        check(lib.cdk_server_stats(self.h, len(keys), _dp(s)))
        return dict(zip(keys, s.tolist()))

    def close(self):
        if self.h:
            check(lib.cdk_server_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- host-side input producers ----------------------------------------
class Rng:
    """std::mt19937_64 as the reference's drivers use (NoMini:203-204)."""

    def __init__(self, seed):
        h = C.c_void_p()
        check(lib.cdk_rng_create(C.c_uint64(seed), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            check(lib.cdk_rng_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def QuaternionUniformSampling(rng, n):
    """QuaternionUniformSampling.cpp:5-24 -> [n][4] (w, x, y, z)."""
    out = np.empty((int(n), 4), np.float32)
    check(lib.cdk_quaternion_uniform_sampling(rng.h, int(n), _fp(out)))
    return out


def GetIdxOfAtomsForVdwRadius(atom_radii, grid_radii):
    """GetIdxOfAtomsForVdwRadius.cpp:4-35 -> (num_used, idx_used, atom_radius_index)."""
    ar = _f32(atom_radii)
    gr = _f32(grid_radii)
    nu = C.c_int()
    used = np.zeros(gr.shape[0], np.int32)
    ari = np.zeros(ar.shape[0], np.int32)
    check(lib.cdk_idx_of_atoms_for_vdw_radius(ar.shape[0], _fp(ar), gr.shape[0], _fp(gr), C.byref(nu), _ip(used),
                                              _ip(ari)))
    return nu.value, used[: nu.value].copy(), ari


def sort_index(values):
    """Energy order with ties by position (stable restatement of sort_index,
    GeneDiverseConformations.h:9-20)."""
    return np.argsort(np.asarray(values), kind="stable")


def synth_protein_atoms(seed, natom, box):
    """(xyz, charge, epsilon, radius) of the synthetic protein, float64."""
    x = np.empty((natom, 3), np.float64)
    q = np.empty(natom, np.float64)
    e = np.empty(natom, np.float64)
    r = np.empty(natom, np.float64)
    check(lib.cdk_synth_protein_atoms(C.c_uint64(seed), int(natom), float(box), _dp(x), _dp(q), _dp(e), _dp(r)),
          "cdk_synth_protein_atoms")
    return x, q, e, r


def synth_protein(seed, natom, box, n, spacing):
    v = np.empty((9, n, n, n), np.float32)
    r = np.empty(8, np.float32)
    mid = np.empty(3, np.float32)
    ln = np.empty(3, np.float32)
    check(lib.cdk_synth_protein(C.c_uint64(seed), int(natom), float(box), int(n), float(spacing), _fp(v), _fp(r),
                                _fp(mid), _fp(ln)), "cdk_synth_protein")
    return v, r, mid, ln


def synth_ligand(seed, natom):
    x = np.empty((natom, 3), np.float32)
    q = np.empty(natom, np.float32)
    e = np.empty(natom, np.float32)
    r = np.empty(natom, np.float32)
    check(lib.cdk_synth_ligand(C.c_uint64(seed), int(natom), _fp(x), _fp(q), _fp(e), _fp(r)))
    return x, q, e, r


def synth_conformers(seed, coords, nconf):
    x = _f32(coords, (-1, 3))
    out = np.empty((int(nconf), x.shape[0], 3), np.float32)
    check(lib.cdk_synth_conformers(C.c_uint64(seed), x.shape[0], _fp(x), int(nconf), _fp(out)))
    return out


def read_grids(path):
    """ReadGrids.cpp:6-68 -> (values [G][X][Y][Z], radii, mid, len, spacing, restraint_k)."""
    ng = C.c_int()
    dims = np.zeros(3, np.int32)
    mid = np.zeros(3, np.float32)
    ln = np.zeros(3, np.float32)
    sp = C.c_float()
    rk = C.c_float()
    p = str(path).encode()
    check(lib.cdk_read_grid_header(p, C.byref(ng), _ip(dims), _fp(mid), _fp(ln), C.byref(sp), C.byref(rk)))
    radii = np.zeros(max(ng.value - 1, 1), np.float32)
    vals = np.zeros((ng.value, dims[0], dims[1], dims[2]), np.float32)
    check(lib.cdk_read_grids(p, _fp(radii), _fp(vals)))
    return vals, radii[: ng.value - 1], mid, ln, sp.value, rk.value


def write_grids(path, values, radii, mid, length, spacing, restraint_k=0.0):
    v = _f32(values)
    dims = _i32(v.shape[1:])
    check(lib.cdk_write_grids(str(path).encode(), v.shape[0], _ip(dims), _fp(_f32(mid)), _fp(_f32(length)),
                              float(spacing), float(restraint_k), _fp(_f32(radii)), _fp(v)))
