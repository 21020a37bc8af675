"""baorec_hip: MI355X (gfx950) reconstruction hot path of BAOrec.jl.

The compute runs in libbaorec_hip.so (hand-written HIP); this package is the
Python mirror of the Julia binding that drives its C ABI.
"""
from ._lib import LIB_PATH, load  # noqa: F401
from .api import (  # noqa: F401
    BaorecError,
    BoundsError,
    Context,
    Cosmology,
    IterativeRecon,
    MultigridRecon,
    ResultCache,
    cartesian_to_sky,
    cic_,
    comoving_distance,
    default_context,
    fmg,
    irfft3d,
    jacobi_,
    k_vec,
    power_multipoles,
    prolong_,
    read_catalog,
    read_cic_,
    read_shifts,
    reconstructed_overdensity_,
    reconstructed_positions,
    run_batch,
    autotune,
    tune_report,
    tune_save,
    tune_load,
    reconstructed_potential_,
    reduce_,
    residual_,
    rfft3d,
    run_,
    setup_box,
    setup_overdensity_,
    sky_to_cartesian,
    smooth_,
    vcycle_,
    write_npy,
    x_vec,
)

__version__ = "0.1.0"
