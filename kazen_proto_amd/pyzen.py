"""`import kazen_proto_amd.pyzen as pyzen`: the reference module's layout
(src/binding/binding.cpp:12-20): pyzen.api (scene, renderer, callbacks),
pyzen.vec and pyzen.mat (vector and matrix types)."""
from . import pyzen_api as api  # noqa: F401
from .pyzen_types import mat, vec  # noqa: F401
