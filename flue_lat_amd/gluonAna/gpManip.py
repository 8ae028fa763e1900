"""Mirror of python/FLUE/gluonAna/gpManip.py: propagator post-processing on the host,
through the library's C entry points (the reference runs them on the host too)."""
from ..compiled import flue_c as fc


def q_average(NQ, q, D, q_slices, DOUT=None, QOUT=None, Qcount=None):
    """python/FLUE/gluonAna/gpManip.py:6-7 -> (DOUT, QOUT, Qcount).  The output arguments
    of the reference's signature are accepted and ignored: the results are returned."""
    return fc.q_average_c(NQ, q, D, q_slices)


def cone_cut(NQ, radius, Q, D, D4, NT, NS, angleIN, xiIN, IRCutIN, IRRadiusIN):
    """python/FLUE/gluonAna/gpManip.py:9-13 -> (QOUT, Dout, D4out, qcount)."""
    return fc.cone_cut_c(NQ, radius, Q, D, D4, NT, NS, angleIN, xiIN, IRCutIN, IRRadiusIN)
