from .simpli import Simpli

__all__ = ['Simpli']
