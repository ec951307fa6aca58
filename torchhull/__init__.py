"""Drop-in import name: ``import torchhull`` resolves to the MI355X implementation (torchhull_amd)."""

from torchhull_amd import *  # noqa: F401,F403
from torchhull_amd import __all__, __version__  # noqa: F401
