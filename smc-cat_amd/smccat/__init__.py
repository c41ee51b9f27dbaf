"""smccat: MI355X-native particle-likelihood engine for Smc-CAT (Python view of the C-ABI)."""
from .engine import Engine, EngineError, ParticleBatch, discretize_gamma, lg_base_matrix  # noqa: F401
