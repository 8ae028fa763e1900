"""Mirror of FLUE.gluonCalc (python/FLUE/gluonCalc/__init__.py:1-4) plus the
gluonic-observable module ``wloops``."""
from .su3matrixops import multiplymatmat
from .gluonprop import calc_mom_space_scalarD, scalarGluonProp
from . import wloops

__all__ = ['multiplymatmat', 'calc_mom_space_scalarD', 'scalarGluonProp', 'wloops']
