"""MI355X (gfx950) FFT translational-rotational pose search.

The compute path is libcdk.so (hand-written HIP kernels + C++ host engine)
reached through its C-ABI (include/cdk.h); this package is a thin binding.
"""
from ._lib import CdkError, ENGINE_FUSED, ENGINE_REFERENCE, LIB_PATH  # noqa: F401
from .api import (  # noqa: F401
    Batch,
    Context,
    GetIdxOfAtomsForVdwRadius,
    Ligand,
    Protein,
    QuaternionUniformSampling,
    Rng,
    Server,
    generate_conformers,
    grid_for_openmm,
    rank_poses,
    refine_poses,
    rescore_poses,
    read_grids,
    search,
    sort_index,
    synth_conformers,
    synth_ligand,
    synth_protein,
    synth_protein_atoms,
    write_grids,
    write_poses,
)
