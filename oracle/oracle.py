"""ORACLE - test infrastructure only (tests/, smoke check, bench CPU baseline).

ctypes wrappers over liboracle.so (oracle.cpp, a CPU restatement of the
reference's search path) and, when it has been built, _ref/libcdkref.so (the
reference's own pure-math sources compiled by oracle/Makefile).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# CDK_ORACLE_LIBRARY: the sanitizer build of the oracle (CPU test runs)
ORACLE_PATH = os.environ.get("CDK_ORACLE_LIBRARY") or os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libcdkref.so")

FP = C.POINTER(C.c_float)
DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)


def _fp(a):
    return a.ctypes.data_as(FP)


def _ip(a):
    return a.ctypes.data_as(IP)


def _dp(a):
    return a.ctypes.data_as(DP)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


_orc = C.CDLL(ORACLE_PATH)
_orc.orc_sparse_energy_f64.restype = C.c_double
_orc.orc_sparse_energy_f64.argtypes = [FP, C.c_int, C.c_float, C.c_int, FP, FP, FP, FP, C.c_int, IP, C.c_int,
                                       C.c_int, C.c_int]
_orc.orc_filter_quaternions.restype = C.c_int
_orc.orc_search_pose.restype = C.c_int
_orc.orc_pose_box.restype = C.c_int
_orc.orc_quaternion_uniform_sampling.argtypes = [C.c_uint64, C.c_int, C.c_int, FP]

_ref_lib = None
_cpubase_lib = None
CPUBASE_PATH = os.path.join(HERE, "libcpubase.so")
CPUBASE_FLAGS = "g++ -O3 -march=x86-64-v3 -ffp-contract=off"


def correlate_window_full_timed(Pused, L, win, threads):
    """correlate_window_full from libcpubase.so (the same source built -O3
    for AVX2/FMA hosts): the timed CPU baseline."""
    global _cpubase_lib
    if _cpubase_lib is None:
        _cpubase_lib = C.CDLL(CPUBASE_PATH)
    _cpubase_lib.orc_set_threads(int(threads))
    P = _f32(Pused)
    L = _f32(L)
    Cn, bx, by, bz = L.shape
    N = P.shape[-1]
    wx, wy, wz = [int(w) for w in win]
    E = np.empty((wx, wy, wz), np.float32)
    _cpubase_lib.orc_correlate_window_full_f32(_fp(P), N, _fp(L), Cn, bx, by, bz, wx, wy, wz, _fp(E))
    return E


def _ref():
    """The compiled reference subset, loaded on first use only (CPU tests
    and the golden-vector script ask for it; nothing else maps it)."""
    global _ref_lib
    if _ref_lib is None:
        if not os.path.exists(REF_PATH):
            raise RuntimeError("oracle/_ref/libcdkref.so missing: run `make -C oracle ref`")
        lib = C.CDLL(REF_PATH)
        lib.ref_filter_quaternions.restype = C.c_int
        lib.ref_quaternion_uniform_sampling.argtypes = [C.c_uint64, C.c_int, C.c_int, FP]
        lib.ref_read_grids.restype = C.c_int
        lib.ref_read_grids.argtypes = [C.c_char_p, IP, IP, FP, FP, FP, FP, FP, C.c_int, FP, C.c_longlong]
        _ref_lib = lib
    return _ref_lib


def have_ref():
    return os.path.exists(REF_PATH)


# ---- restated functions --------------------------------------------------
def rotate(quats, coords, lib=None):
    q = _f32(quats).reshape(-1, 4)
    x = _f32(coords).reshape(-1, 3)
    out = np.empty((q.shape[0], x.shape[0], 3), np.float32)
    if lib == "ref":
        _ref().ref_rotate_batch(q.shape[0], x.shape[0], _fp(q), _fp(x), _fp(out))
    else:
        _orc.orc_rotate_batch(q.shape[0], x.shape[0], _fp(q), _fp(x), _fp(out))
    return out


def min_max_coors(rot, lib=None):
    r = _f32(rot)
    nq, na = r.shape[0], r.shape[1]
    mn = np.empty((nq, 3), np.float32)
    mx = np.empty((nq, 3), np.float32)
    if lib == "ref":
        _ref().ref_min_coors(nq, na, _fp(r), _fp(mn))
        _ref().ref_max_coors(nq, na, _fp(r), _fp(mx))
    else:
        _orc.orc_min_coors(nq, na, _fp(r), _fp(mn))
        _orc.orc_max_coors(nq, na, _fp(r), _fp(mx))
    return mn, mx


def idx_of_atoms_for_vdw_radius(atom_radii, grid_radii, lib=None):
    ar = _f32(atom_radii)
    gr = _f32(grid_radii)
    nu = C.c_int()
    used = np.zeros(gr.shape[0], np.int32)
    ari = np.zeros(ar.shape[0], np.int32)
    f = _ref().ref_idx_of_atoms_for_vdw_radius if lib == "ref" else _orc.orc_idx_of_atoms_for_vdw_radius
    f(ar.shape[0], _fp(ar), gr.shape[0], _fp(gr), C.byref(nu), _ip(used), _ip(ari))
    return nu.value, used[: nu.value].copy(), ari


def csr_of_radius(atom_radius_index, nv):
    offs = np.zeros(nv + 1, np.int32)
    atoms = []
    for r in range(nv):
        offs[r] = len(atoms)
        atoms.extend(np.nonzero(np.asarray(atom_radius_index) == r)[0].tolist())
    offs[nv] = len(atoms)
    return offs, _i32(atoms if atoms else [0])


def fill_ligand_grid(coors, mincoors, charge, eps, num_used, idx_used, atom_radius_index, nv, dims, spacing,
                     grid=None, lib=None):
    co = _f32(coors)
    mc = _f32(mincoors).reshape(-1, 3)
    nq = mc.shape[0]
    q = _f32(charge)
    e = _f32(eps)
    na = q.shape[0]
    offs, atoms = csr_of_radius(atom_radius_index, nv)
    used = _i32(idx_used if len(idx_used) else [0])
    X, Y, Z = [int(v) for v in dims]
    if grid is None:
        grid = np.zeros((nq, num_used + 1, X, Y, Z), np.float32)
    g = np.ascontiguousarray(grid, dtype=np.float32)
    if lib == "ref":
        _ref().ref_fill_ligand_grid(nq, na, _fp(co), _fp(mc), _fp(q), _fp(e), num_used, _ip(used), nv, _ip(offs),
                                  _ip(atoms), X, Y, Z, C.c_float(spacing), _fp(g))
    else:
        _orc.orc_fill_ligand_grid(nq, na, _fp(co), _fp(mc), _fp(q), _fp(e), num_used, _ip(used), _ip(offs),
                                  _ip(atoms), X, Y, Z, C.c_float(spacing), _fp(g))
    return g


def quaternion_uniform_sampling(seed, ncall, n, lib=None):
    out = np.empty((ncall, n, 4), np.float32)
    f = _ref().ref_quaternion_uniform_sampling if lib == "ref" else _orc.orc_quaternion_uniform_sampling
    f(C.c_uint64(seed), ncall, n, _fp(out))
    return out


def filter_quaternions(coor, quats, lens, lib=None):
    x = _f32(coor).reshape(-1, 3)
    q = _f32(quats).reshape(-1, 4)
    if lib == "ref":
        out = np.empty_like(q)
        n = _ref().ref_filter_quaternions(_fp(x), x.shape[0], q.shape[0], _fp(q), C.c_float(lens[0]),
                                        C.c_float(lens[1]), C.c_float(lens[2]), q.shape[0], _fp(out))
        return out[:n].copy()
    kept = np.zeros(q.shape[0], np.int32)
    n = _orc.orc_filter_quaternions(_fp(x), x.shape[0], q.shape[0], _fp(q), C.c_float(lens[0]), C.c_float(lens[1]),
                                    C.c_float(lens[2]), _ip(kept))
    return q[kept[:n]].copy()


def read_grids_ref(path, max_values=1 << 22):
    """ReadGrids.cpp:6-68, the compiled reference reader: (values [G][X][Y][Z],
    radii, mid, len, spacing, restraint_k)."""
    ng = C.c_int()
    dims = np.zeros(3, np.int32)
    mid = np.zeros(3, np.float32)
    ln = np.zeros(3, np.float32)
    sp = C.c_float()
    rk = C.c_float()
    radii = np.zeros(64, np.float32)
    vals = np.zeros(max_values, np.float32)
    rc = _ref().ref_read_grids(str(path).encode(), C.byref(ng), _ip(dims), _fp(mid), _fp(ln), C.byref(sp),
                               C.byref(rk), _fp(radii), 64, _fp(vals), max_values)
    if rc != 0:
        raise ValueError("grid file larger than the buffers")
    n = int(ng.value) * int(np.prod(dims))
    return (vals[:n].reshape(ng.value, *dims).copy(), radii[: ng.value - 1].copy(), mid, ln, sp.value, rk.value)


def argmin_first(E):
    E = _f32(E)
    wx, wy, wz = E.shape
    e = C.c_float()
    idx = np.zeros(3, np.int32)
    _orc.orc_argmin_first(_fp(E), wx, wy, wz, C.byref(e), _ip(idx))
    return e.value, idx


def set_threads(n):
    _orc.orc_set_threads(int(n))


def correlate_window_full(Pused, L, win):
    """The literal cpuNaive loop (every box cell), FP32."""
    P = _f32(Pused)
    L = _f32(L)
    Cn, bx, by, bz = L.shape
    N = P.shape[-1]
    wx, wy, wz = [int(w) for w in win]
    E = np.empty((wx, wy, wz), np.float32)
    _orc.orc_correlate_window_full_f32(_fp(P), N, _fp(L), Cn, bx, by, bz, wx, wy, wz, _fp(E))
    return E


def correlate_window(Pused, L, win, double=False):
    """Dense direct correlation (cpuNaive.cpp:184-216) of a ligand box L
    [C][bx][by][bz] with protein channels Pused [C][N][N][N] over win."""
    P = _f32(Pused)
    L = _f32(L)
    Cn, bx, by, bz = L.shape
    N = P.shape[-1]
    wx, wy, wz = [int(w) for w in win]
    if double:
        E = np.empty((wx, wy, wz), np.float64)
        _orc.orc_correlate_window_f64(_fp(P), N, _fp(L), Cn, bx, by, bz, wx, wy, wz, _dp(E))
    else:
        E = np.empty((wx, wy, wz), np.float32)
        _orc.orc_correlate_window_f32(_fp(P), N, _fp(L), Cn, bx, by, bz, wx, wy, wz, _fp(E))
    return E


class PoseProblem:
    """One protein + one ligand, the oracle's view of a search."""

    def __init__(self, values, grid_radii, spacing, charge, eps, radius):
        self.values = _f32(values)
        self.N = self.values.shape[-1]
        self.h = float(spacing)
        self.charge = _f32(charge)
        self.eps = _f32(eps)
        self.radius = _f32(radius)
        self.natom = self.charge.shape[0]
        nv = len(grid_radii)
        self.nv = nv
        self.num_used, self.idx_used, self.ari = idx_of_atoms_for_vdw_radius(radius, grid_radii)
        # used protein channels: vdW in used order, elec last (NoMini:136-146)
        self.Pused = np.ascontiguousarray(
            np.concatenate([self.values[self.idx_used], self.values[nv:nv + 1]], axis=0))
        ch = -np.ones(self.natom, np.int32)
        for c, r in enumerate(self.idx_used):
            ch[self.ari == r] = c
        self.atom_channel = ch

    def search_pose(self, coords, q, double=False, want_E=False):
        x = _f32(coords).reshape(-1, 3)
        q = _f32(q).reshape(4)
        emin = C.c_float()
        idx = np.zeros(3, np.int32)
        minc = np.zeros(3, np.float32)
        win = np.zeros(3, np.int32)
        used = _i32(self.idx_used if len(self.idx_used) else [0])
        Ebuf = None
        Eptr = None
        if want_E:
            Ebuf = np.empty(self.N ** 3, np.float64)
            Eptr = _dp(Ebuf)
        ok = _orc.orc_search_pose(_fp(self.Pused), self.N, C.c_float(self.h), self.natom, _fp(x), _fp(q),
                                  _fp(self.charge), _fp(self.eps), self.num_used, _ip(used), _ip(self.ari),
                                  int(double), C.byref(emin), _ip(idx), _fp(minc), _ip(win), Eptr)
        E = None
        if want_E and ok:
            E = Ebuf[: int(np.prod(win))].reshape(tuple(win))
        return dict(energy=emin.value, index=idx, mincoor=minc, window=win, valid=bool(ok), E=E)

    def pose_box(self, coords, q):
        x = _f32(coords).reshape(-1, 3)
        q = _f32(q).reshape(4)
        box = np.zeros(3, np.int32)
        win = np.zeros(3, np.int32)
        minc = np.zeros(3, np.float32)
        used = _i32(self.idx_used if len(self.idx_used) else [0])
        n = _orc.orc_pose_box(self.N, C.c_float(self.h), self.natom, _fp(x), _fp(q), _fp(self.charge),
                              _fp(self.eps), self.num_used, _ip(used), _ip(self.ari), _ip(box), _ip(win), _fp(minc),
                              None, C.c_size_t(0))
        L = np.zeros(n, np.float32)
        _orc.orc_pose_box(self.N, C.c_float(self.h), self.natom, _fp(x), _fp(q), _fp(self.charge), _fp(self.eps),
                          self.num_used, _ip(used), _ip(self.ari), _ip(box), _ip(win), _fp(minc), _fp(L),
                          C.c_size_t(n))
        return L.reshape(self.num_used + 1, *box), box, win, minc

    def sparse_search(self, coords, conf, quats, t_at=None, full=True):
        """orc_sparse_search_f64: full-window FP64 argmin of every pose
        (conformer coords[conf[p]] rotated by quats[p]); with t_at [n][3], E64
        and the absolute scale S at those translations too (full=False: only
        those)."""
        x = _f32(coords).reshape(-1, self.natom, 3)
        cf = _i32(conf).reshape(-1)
        q = _f32(quats).reshape(-1, 4)
        n = q.shape[0]
        assert cf.shape[0] == n and (cf.max(initial=0) < x.shape[0])
        valid = np.zeros(n, np.int32)
        emin = np.zeros(n, np.float64)
        idx = np.zeros((n, 3), np.int32)
        e_at = np.full(n, np.nan)
        s_at = np.full(n, np.nan)
        ta = None if t_at is None else _i32(t_at).reshape(n, 3)
        _orc.orc_sparse_search_f64(_fp(self.Pused), self.N, C.c_float(self.h), self.natom, _fp(x), _ip(cf), _fp(q), n,
                                   _fp(self.charge), _fp(self.eps), self.num_used, _ip(self.atom_channel),
                                   None if ta is None else _ip(ta), int(bool(full)), _ip(valid), _dp(emin),
                                   _ip(idx), _dp(e_at),
                                   _dp(s_at))
        return dict(valid=valid.astype(bool), emin=emin, index=idx, e_at=e_at, s_at=s_at)

    def sparse_energy(self, rot, minc, t):
        r = _f32(rot).reshape(-1, 3)
        m = _f32(minc).reshape(3)
        return _orc.orc_sparse_energy_f64(_fp(self.Pused), self.N, C.c_float(self.h), self.natom, _fp(r), _fp(m),
                                          _fp(self.charge), _fp(self.eps), self.num_used, _ip(self.atom_channel),
                                          int(t[0]), int(t[1]), int(t[2]))


def build_grids(xyz, charge, epsilon, radius, probe_radii, n, gmin, h, rc_lj=8.0, rc_el=12.0, lj_cap=3.0,
                el_cap=20.0):
    """Protein potential grids from atoms, brute force FP64 (orc_build_grids)."""
    x = np.asarray(xyz, np.float64).reshape(-1, 3)
    atoms = np.ascontiguousarray(np.concatenate(
        [x, np.asarray(charge, np.float64)[:, None], np.asarray(radius, np.float64)[:, None],
         np.asarray(epsilon, np.float64)[:, None]], axis=1))
    pr = np.ascontiguousarray(np.asarray(probe_radii, np.float32).astype(np.float64))
    g = np.ascontiguousarray(np.asarray(gmin, np.float64))
    out = np.empty((pr.shape[0] + 1, n, n, n), np.float32)
    _orc.orc_build_grids(atoms.shape[0], _dp(atoms), pr.shape[0], _dp(pr), int(n), _dp(g), C.c_double(h),
                         C.c_double(rc_lj), C.c_double(rc_el), C.c_double(lj_cap), C.c_double(el_cap), _fp(out))
    return out


def chain_rotors(natom):
    """Rotors, moving masks and 1-2/1-3/1-4 exclusions of a linear chain
    (bond i-(i+1)): rotor (i-1, i, i+1, i+2) turns atoms i+2.. about i->i+1."""
    rot = np.array([[i - 1, i, i + 1, i + 2] for i in range(1, natom - 2)], np.int32).reshape(-1, 4)
    mv = np.zeros((rot.shape[0], natom), np.uint8)
    for r, (a, b, c, d) in enumerate(rot):
        mv[r, c + 1:] = 1
    idx = np.arange(natom)
    ex = (np.abs(idx[:, None] - idx[None, :]) <= 3).astype(np.uint8)
    return rot, mv, ex


def conformers(coords, rotors, moving, exclude, ntrial, seed, clash_dist):
    """orc_conformers: trial coordinates [ntrial][natom][3] and keep flags."""
    x = np.ascontiguousarray(np.asarray(coords, np.float64).reshape(-1, 3))
    na = x.shape[0]
    r = np.ascontiguousarray(np.asarray(rotors, np.int32).reshape(-1, 4))
    mv = np.ascontiguousarray(np.asarray(moving, np.uint8).reshape(r.shape[0], na))
    ex = np.ascontiguousarray(np.asarray(exclude, np.uint8).reshape(na, na))
    out = np.empty((int(ntrial), na, 3), np.float64)
    ok = np.empty(int(ntrial), np.int32)
    _orc.orc_conformers(na, _dp(x), r.shape[0], _ip(r), mv.ctypes.data_as(C.POINTER(C.c_uint8)),
                        ex.ctypes.data_as(C.POINTER(C.c_uint8)), int(ntrial), C.c_uint64(seed),
                        C.c_double(clash_dist), _dp(out), _ip(ok))
    return out, ok.astype(bool)
