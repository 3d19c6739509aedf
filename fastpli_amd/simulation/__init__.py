from .autograd import simulate_stack
from .serving import SimulationServer
from .simpli import MarchFields, Simpli

__all__ = ['MarchFields', 'Simpli', 'SimulationServer', 'simulate_stack']
