"""ctypes binding of libcdk.so (include/cdk.h).

The shared library is the product: every compute call goes through it.  If it
is missing or fails to load, importing this module raises; there is no
fallback path.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# CDK_LIBRARY: another build of the same library (the host-sanitizer
# variant the CPU tests can run under); default the in-tree build.
LIB_PATH = os.environ.get("CDK_LIBRARY") or os.path.join(_HERE, "libcdk.so")

OK = 0
ERR_INVALID = 1
ERR_HIP = 2
ERR_NOMEM = 3
ERR_UNSUPPORTED = 4
ERR_PENDING = 5
ERR_BUDGET = 6
ERR_CANCELLED = 7

ENGINE_REFERENCE = 0
ENGINE_FUSED = 1

_STATUS_NAMES = {
    ERR_INVALID: "CDK_ERR_INVALID",
    ERR_HIP: "CDK_ERR_HIP",
    ERR_NOMEM: "CDK_ERR_NOMEM",
    ERR_UNSUPPORTED: "CDK_ERR_UNSUPPORTED",
    ERR_PENDING: "CDK_ERR_PENDING",
    ERR_BUDGET: "CDK_ERR_BUDGET",
    ERR_CANCELLED: "CDK_ERR_CANCELLED",
}


class CdkError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (_STATUS_NAMES.get(code, str(code)), msg))
        self.code = code


if not os.path.exists(LIB_PATH):
    raise ImportError(
        "libcdk.so not found at %s; build it with `python -c 'import __graft_entry__ as g; g.build()'` "
        "or `make -C cdocker_amd/csrc`" % LIB_PATH)

lib = C.CDLL(LIB_PATH)

P = C.c_void_p
PP = C.POINTER(C.c_void_p)
I = C.c_int
I64 = C.c_int64
U64 = C.c_uint64
F = C.c_float
FP = C.POINTER(C.c_float)
DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)
I64P = C.POINTER(C.c_int64)
CP = C.c_char_p

# name: argtypes (all return int status unless listed in _RET)
_SIGS = {
    "cdk_version": [],
    "cdk_context_create": [I, PP],
    "cdk_context_destroy": [P],
    "cdk_context_set_option": [P, CP, I64],
    "cdk_context_get_option": [P, CP, I64P],
    "cdk_context_stage_times": [P, DP, IP, I],
    "cdk_protein_create": [P, I, I, I, I, FP, FP, F, FP, FP, PP],
    "cdk_protein_build": [P, I, DP, DP, DP, DP, I, FP, I, F, FP, FP, C.c_double, C.c_double, C.c_double,
                          C.c_double, PP],
    "cdk_protein_grids": [P, FP],
    "cdk_protein_destroy": [P],
    "cdk_protein_spectrum": [P, I, FP],
    "cdk_protein_bytes": [P, I64P],
    "cdk_ligand_create": [P, I, FP, FP, FP, PP],
    "cdk_ligand_destroy": [P],
    "cdk_ligand_channels": [P, IP, IP, IP],
    "cdk_search": [P, P, P, I, FP, I, FP, I, I, FP, IP, FP],
    "cdk_rank_poses": [P, I, I, FP, FP, IP, FP, I, IP, DP],
    "cdk_rescore_poses": [P, P, I, DP, DP, DP, IP],
    "cdk_refine_poses": [P, P, I, DP, I, C.c_double, DP],
    "cdk_write_poses": [CP, I, I, I, DP, FP, C.POINTER(CP), C.POINTER(CP), CP],
    "cdk_grid_for_openmm": [P, I, DP, DP],
    "cdk_generate_conformers": [P, I, DP, I, IP, C.POINTER(C.c_uint8), C.POINTER(C.c_uint8), I, U64, C.c_double, I,
                                DP, IP, IP],
    "cdk_batch_create": [P, P, PP],
    "cdk_batch_add": [P, P, I, FP, I, FP, I, IP],
    "cdk_batch_run": [P, I],
    "cdk_batch_poses": [P, I64P],
    "cdk_batch_results": [P, I, FP, IP, FP],
    "cdk_batch_geometry": [P, I, IP, IP],
    "cdk_batch_destroy": [P],
    "cdk_server_create": [P, I, PP, I64, I, PP],
    "cdk_server_submit": [P, I, I, FP, FP, FP, I, FP, I, FP, I, I64P],
    "cdk_server_result": [P, I64, I, FP, IP, FP, DP],
    "cdk_server_stats": [P, I, DP],
    "cdk_server_destroy": [P],
    "cdk_rotate": [P, I, I, FP, FP, FP],
    "cdk_rotate_aabb": [P, I, I, FP, FP, I, F, FP, FP, FP, IP],
    "cdk_fill_ligand_grid": [P, I, I, FP, FP, FP, FP, I, IP, I, IP, IP, I, I, I, F, FP],
    "cdk_fft_r2c_3d": [P, I, I, FP, FP],
    "cdk_fft_c2r_3d": [P, I, I, FP, FP],
    "cdk_conj_mult": [P, I, I, I64, FP, FP],
    "cdk_sum_grids": [P, I, I, I64, I64, FP, FP],
    "cdk_argmin_window": [P, I, I, FP, IP, FP, IP],
    "cdk_rank_poses_quats": [P, I, I, FP, FP, FP, IP, FP, I, IP, DP],
    "cdk_rotate_host": [I, I, FP, FP, FP],
    "cdk_min_max_coors": [I, I, FP, FP, FP],
    "cdk_quaternion_uniform_sampling_mt": [P, I, FP],
    "cdk_argmin_window_noalloc": [P, I, I, P, P, P, P, P],
    "cdk_rng_create": [U64, PP],
    "cdk_rng_destroy": [P],
    "cdk_quaternion_uniform_sampling": [P, I, FP],
    "cdk_filter_quaternions": [P, FP, I, I, FP, F, F, F, I, U64, IP, FP],
    "cdk_idx_of_atoms_for_vdw_radius": [I, FP, I, FP, IP, IP, IP],
    "cdk_synth_protein_atoms": [U64, I, F, DP, DP, DP, DP],
    "cdk_synth_protein": [U64, I, F, I, F, FP, FP, FP, FP],
    "cdk_synth_ligand": [U64, I, FP, FP, FP, FP],
    "cdk_synth_conformers": [U64, I, FP, I, FP],
    "cdk_read_grid_header": [CP, IP, IP, FP, FP, FP, FP],
    "cdk_read_grids": [CP, FP, FP],
    "cdk_write_grids": [CP, I, IP, FP, FP, F, F, FP, FP],
}

lib.cdk_last_error.restype = C.c_char_p
lib.cdk_last_error.argtypes = []
for _name, _args in _SIGS.items():
    _fn = getattr(lib, _name)
    _fn.argtypes = _args
    _fn.restype = C.c_int

DECLARED = sorted(list(_SIGS) + ["cdk_last_error"])


def last_error():
    return lib.cdk_last_error().decode("utf-8", "replace")


def check(rc, what=""):
    if rc != OK:
        raise CdkError(rc, (what + ": " if what else "") + last_error())
    return rc
