"""smccat: MI355X-native particle-likelihood engine for Smc-CAT (Python view of the C-ABI)."""
from .engine import (Engine, EngineError, ParticleBatch, ProposalNetwork, discretize_gamma,  # noqa: F401
                     lg_base_matrix)
