from .autograd import simulate_stack
from .serving import SimulationServer
from .simpli import Simpli

__all__ = ['Simpli', 'SimulationServer', 'simulate_stack']
