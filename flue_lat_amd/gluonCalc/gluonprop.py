"""Mirror of python/FLUE/gluonCalc/gluonprop.py: the scalar gluon propagator, computed on
the GPU through libflue_hip.so.

``calc_mom_space_scalarD`` takes the Fourier-transformed field (numpy.fft.fftn of the
gauge field over its four lattice axes), ``scalarGluonProp`` the position-space field.
``GaugeField.gluon_propagator`` does the transform and the momentum-space step on the
device for a field already held there.
"""
from ..compiled import flue_c as fc


def calc_mom_space_scalarD(U, mu_start, xi):
    """python/FLUE/gluonCalc/gluonprop.py:3-4"""
    return fc.calc_mom_space_scalard_c(U, mu_start, xi)


def scalarGluonProp(U, xi):
    """python/FLUE/gluonCalc/gluonprop.py:7-9"""
    return fc.scalargluonprop_c(U, xi)
