"""Jackknife resampling on the host (src/jack.f90:1-63).

Complement(data) is the ensemble mean; Complement(data, k) for k = 1, 2, 3 removed
configurations gives the complement means of every first-order subset, optionally with
the configurations icon, jcon (0-based here) removed as well, as the reference's
Complement1..3.  jackknife(c) takes c[0] = the mean and c[1:] = first-order complement
means and returns (error, bias) as JackKnife_wp.
"""
import numpy as np


def Complement(data, icon=None, jcon=None, order=None):
    """Complement means (src/jack.f90:18-46).

    order 0: the mean of data (Complement0).
    order 1: (sum - data) / (n - 1) for every configuration (Complement1).
    order 2: (sum - data[icon] - data) / (n - 2) (Complement2).
    order 3: (sum - data[icon] - data[jcon] - data) / (n - 3) (Complement3).
    With order None it is 0 when neither icon nor jcon is given, else 2 or 3."""
    d = np.asarray(data, dtype=np.float64)
    n = d.shape[0]
    if order is None:
        order = 0 if icon is None else (2 if jcon is None else 3)
    s = d.sum(axis=0)
    if order == 0:
        return s / n
    if order == 1:
        return (s - d) / (n - 1)
    if order == 2:
        return (s - d[icon] - d) / (n - 2)
    if order == 3:
        return (s - d[icon] - d[jcon] - d) / (n - 3)
    raise ValueError("order must be 0..3")


def jackknife(c):
    """(err, bias) from c[0] = mean, c[1:] = first-order complement means
    (JackKnife_wp, src/jack.f90:48-59)."""
    c = np.asarray(c, dtype=np.float64)
    n = c.shape[0] - 1
    avg = c[1:].sum(axis=0) / n
    err = np.sqrt(((c[1:] - avg) ** 2).sum(axis=0) * (n - 1) / n)
    return err, c[0] - avg


def jackknife_mean_err(data):
    """(mean, error) of data by first-order jackknife."""
    c = np.concatenate([[Complement(data)], Complement(data, order=1)])
    err, _ = jackknife(c)
    return float(c[0]), float(err)
