"""Propagator post-processing (python/FLUE/gluonAna/gpManip.py) is outside the scope of
flue_lat_amd; the names raise NotImplementedError."""
from ..compiled import flue_c as fc


def q_average(NQ, q, D, q_slices, DOUT=None, QOUT=None, Qcount=None):
    return fc.q_average_c(NQ, q, D, q_slices)


def cone_cut(NQ, radius, Q, D, D4, NT, NS, angleIN, xiIN, IRCutIN, IRRadiusIN):
    return fc.cone_cut_c(NQ, radius, Q, D, D4, NT, NS, angleIN, xiIN, IRCutIN, IRRadiusIN)
