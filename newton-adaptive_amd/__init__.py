"""MI355X-native batched step-doubling rigid-body step with Newton's Model/State/Control/
Contacts API (mardigiorgio/newton-adaptive ``newton/__init__.py:77-93``,
``newton/solvers.py:409-462``).

Import with ``importlib.import_module("newton-adaptive_amd")`` (the directory name is not a
Python identifier).
"""

from .builder import ModelBuilder
from .contacts import CollisionPipeline, Contacts
from .model import Control, GeoType, Model, State, eval_fk
from .solvers import SolverMuJoCoAdaptive, SolverSAPAdaptive
from .serving import RolloutRequest, ServingEngine, requests_from_model, world_arrays
from . import mujoco_adapter, rl_backend
from . import scenes


class solvers:  # noqa: N801 - namespace mirroring ``newton.solvers``
    SolverSAPAdaptive = SolverSAPAdaptive
    SolverMuJoCoAdaptive = SolverMuJoCoAdaptive


def build(verbose: bool = False) -> str:
    """Compile the gfx950 library in-tree."""
    from ._lib import build as _build

    return _build(verbose)


__all__ = [
    "ModelBuilder", "Model", "State", "Control", "Contacts", "CollisionPipeline", "GeoType", "eval_fk",
    "SolverSAPAdaptive", "SolverMuJoCoAdaptive", "solvers", "scenes", "build",
]
