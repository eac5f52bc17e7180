"""MI355X (gfx950) WCSPH step of TitSolver behind a C ABI (libtit_hip.so)."""
from ._abi import (FLAVOUR_CFG, FLAVOUR_REF, KERNEL_WENDLAND_C2,  # noqa: F401
                   KERNEL_WENDLAND_C4, LIB_PATH, load)
from .sph import (FluidEquations, HipContext, LearnedStep,  # noqa: F401
                  ParticleArray, SSPRKIntegrator, TitError, Walls, box_walls,
                  cfg_equations, dam_break, dam_break_walls, nn_gemm,
                  ref_equations, seeded_tanks, surge_front,
                  tessellate)
