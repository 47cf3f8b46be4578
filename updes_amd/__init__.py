"""updes_amd: MI355X-native dense RBF-collocation solver behind Updes' API.

``import updes_amd as updes`` gives the names of ``updes/__init__.py``'s
star re-exports that the collocation path uses: clouds, kernels, nodal
operators, field evaluation, assembly and ``pde_solver`` /
``pde_solver_jit``.  All compute runs in HIP kernels of ``librbfx.so``
(``include/rbfx.h``) on the GPU.
"""
from .utils import (distance, multiquadric, inverse_multiquadric, gaussian, polyharmonic, thin_plate,
                    multiquadric_func, inv_multiquadric_func, gaussian_func, polyharmonic_func, thin_plate_func,
                    make_monomial, make_all_monomials, compute_nb_monomials, SteadySol, rbf_spec)
from .cloud import Cloud, SquareCloud
from ._symbolic import NotLowerable
from .operators import (nodal_value, nodal_gradient, nodal_laplacian, nodal_div_grad, value, value_vec,
                        gradient, gradient_vec, laplacian, laplacian_vec, divergence, divergence_vec,
                        compute_coefficients, get_field_coefficients, core_compute_coefficients, gradient_vals,
                        laplacian_vals, duplicate_robin_coeffs, zerofy_periodic_cond,
                        boundary_conditions_func_to_arr, pde_solver, pde_solver_jit, pde_solver_jit_with_bc,
                        pde_multi_solver, interpolate_field, integrate_field, last_solve_stats, dot_vec, dot_mat)
from .assembly import (assemble_A, assemble_Phi, assemble_P, assemble_invert_A, assemble_K, assemble_op_Phi_P,
                       assemble_bd_Phi_P, assemble_B, assemble_q)
from .adjoint import checkpointed_rollout
from .engine import clear_cache

__version__ = "0.1.0"
RBF = None
MAX_DEGREE = None
DIM = 2
FLOAT64 = True
