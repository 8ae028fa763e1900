"""The gluon propagator (python/FLUE/gluonCalc/gluonprop.py) is outside the scope of
flue_lat_amd; these keep the names and raise NotImplementedError."""
from ..compiled import flue_c as fc


def calc_mom_space_scalarD(U, mu_start, xi):
    return fc.calc_mom_space_scalard_c(U, mu_start, xi)


def scalarGluonProp(U, xi):
    return fc.scalargluonprop_c(U, xi)
