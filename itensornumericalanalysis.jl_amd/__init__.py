"""MI355X engine for the quantics tensor-network hot path of ITensorNumericalAnalysis.jl.

Host-side API mirroring the reference package (src/ITensorNumericalAnalysis.jl:20-66):
construction (continuous_siteinds, *_itn builders, operators) runs on the host;
evaluate / integrate / inner / operate / truncate / gradients, and the function
algebra (multiply, add, partial_integrate) run as HIP kernels
in libqtt.so (csrc/), called through its C-ABI (include/qtt.h).  There is no
CPU fallback for those: without the built library they raise.
"""
from .graphs import NamedGraph, named_comb_tree, named_grid, named_path_graph
from .indsnetworkmap import (DigitMapError, IndsNetworkMap, continuous_siteinds,
                             default_dimension_vertices, reduced_indsnetworkmap)
from .network import ITensorNetworkFunction, Topology, TreeOperator
from .builders import (const_itn, cos_itn, cosh_itn, delta_kernel, delta_p, exp_itn, poly_itn, random_itn, scale,
                       sin_itn, sinh_itn)
from .files import load, save
from .loopy import LoopyNetworkFunction, evaluate_bp
from .algebra import (add, map_to_zero_operator, map_to_zeros, multiply, operator_proj,
                      partial_integrate)
from .operators import (add_operators, backward_shift_op, compile_terms, first_derivative_operator,
                        forward_shift_op, fourth_derivative_operator, fredholm_operator,
                        identity_operator, laplacian_operator, point_to_opsum, second_derivative_operator,
                        stencil, third_derivative_operator)
from .api import (Precision, apply_grad, engine_context, eval_grad, evaluate, evaluate_batch, inner, integrate,
                  integrate_batch, on_stream, operate, operate_batch, operate_batch_pipeline, truncate,
                  tune_evaluate)

real_continuous_siteinds = continuous_siteinds

__all__ = [
    "NamedGraph", "named_grid", "named_comb_tree", "named_path_graph", "IndsNetworkMap",
    "continuous_siteinds", "real_continuous_siteinds", "default_dimension_vertices",
    "DigitMapError", "ITensorNetworkFunction", "TreeOperator", "Topology", "const_itn", "exp_itn",
    "sin_itn", "cos_itn", "cosh_itn", "sinh_itn", "random_itn", "poly_itn", "add", "multiply", "scale",
    "partial_integrate", "reduced_indsnetworkmap", "delta_p", "delta_kernel", "operator_proj",
    "map_to_zero_operator", "map_to_zeros", "point_to_opsum", "compile_terms", "save", "load",
    "LoopyNetworkFunction", "evaluate_bp",
    "stencil", "forward_shift_op", "backward_shift_op", "first_derivative_operator",
    "second_derivative_operator", "third_derivative_operator", "fourth_derivative_operator",
    "laplacian_operator", "identity_operator", "fredholm_operator", "add_operators", "evaluate",
    "integrate", "inner", "operate", "truncate", "eval_grad", "Precision", "engine_context",
    "operate_batch", "operate_batch_pipeline", "integrate_batch", "evaluate_batch", "tune_evaluate", "apply_grad",
]
