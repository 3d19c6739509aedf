from .serving import SimulationServer
from .simpli import Simpli

__all__ = ['Simpli', 'SimulationServer']
