"""pypde_amd: MI355X-native spectral-Galerkin RBC2D time step.

Drop-in surface of preiter93/pypde for the RBC2D hot path (bases, fields,
field operations, solver plans, navier.rbc2d) with float64 device tensors and
hand-written gfx950 HIP kernels behind the C-ABI in include/pypde_amd.h.
"""
from ._lib import PypdeError
from .bases import (Base, Chebyshev, ChebDirichlet, ChebDirichletNeumann, ChebNeumann,
                    DirichletC, NeumannC, MetaBase, GalerkinChebyshev, gauss_lobatto)
from .field import Field, FieldBC, MultiField, SpectralSpace, SpectralSpaceBC
from .field_operations import (avg_vol, avg_x, cheby_to_galerkin, conv_term, convective_term,
                               galerkin_to_cheby, grad, interpolate)
from .solver import (PDMA, Integrator, MetaPlan, Plan_fdma, Plan_numpy, Plan_Poisson,
                     Plan_twodma, PlanLHS, PlanRHS, SolverPlan, eigdecomp)
from .templates import (solverplan_hholtz1d, solverplan_hholtz2d_adi, solverplan_poisson1d,
                        solverplan_poisson2d)

__all__ = [
    "PypdeError", "Base", "Chebyshev", "ChebDirichlet", "ChebNeumann", "ChebDirichletNeumann", "DirichletC", "NeumannC",
    "MetaBase", "GalerkinChebyshev", "gauss_lobatto", "Field", "FieldBC", "MultiField",
    "SpectralSpace", "SpectralSpaceBC", "avg_vol", "avg_x", "cheby_to_galerkin", "conv_term",
    "convective_term", "galerkin_to_cheby", "grad", "interpolate", "PDMA", "Integrator",
    "MetaPlan", "Plan_fdma", "Plan_numpy", "Plan_Poisson", "Plan_twodma", "PlanLHS", "PlanRHS",
    "SolverPlan", "eigdecomp", "solverplan_hholtz1d", "solverplan_hholtz2d_adi",
    "solverplan_poisson1d", "solverplan_poisson2d",
]
