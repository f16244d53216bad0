"""qbhip: host mirror of QuantumBilliards.jl's Kress/Beyn hot path on libqbhip.

The compute runs in libqbhip.so (HIP, gfx950) behind the C ABI of
include/qb.h; this package only arranges inputs and outputs the way the
reference's Julia API does (see qbhip.solvers and qbhip.beyn).
"""
from ._lib import (QB_CFIE_KRESS, QB_DLP_KRESS, QB_FP64, QB_MIXED, QBError, LIB_PATH,
                   EXPORTS, lib)
from .context import Context, default_context, pinned_array
from .geometry import (Billiard, BoundaryPointsCFIE, CircularHoleBilliard, evaluate_points,
                       make_circle, make_full_ellipse, make_full_stadium,
                       make_star_polar_boundary, n_kress)
from .solvers import (CFIE_kress, DLP_kress, Geometry, assemble, assemble_with_derivatives,
                      bessel_eval, boundary_matrix_size, boundary_matrix_size_for_solver,
                      chebyshev_params,
                      construct_boundary_matrices, construct_boundary_matrices_with_derivatives)
from .beyn import (B_from_moments, B_from_qr, Window, construct_B_matrix, footprint,
                   probe_block, probe_count, solve, solve_vect, weyl_count)
from .spectrum import (beyn_disks_from_windows, compute_spectrum_beyn, delta_weyl,
                       imag_k_check_EXPERIMENTAL, plan_weyl_windows, residual_and_norm_select,
                       residual_norms)
from .symmetry import (Reflection, Rotation, XReflection, XYReflection, YReflection,
                       build_symmetry_maps, cfie_orbits, dlp_orbits, node_divisor, orbits_for,
                       rotation_characters)
