from ..compiled import flue_c as fc


def multiplymatmat(left, right):
    """python/FLUE/gluonCalc/su3matrixops.py:3-4"""
    return fc.multiplymatmat_c(left, right)
