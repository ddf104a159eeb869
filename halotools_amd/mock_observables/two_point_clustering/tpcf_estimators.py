"""Two-point estimators (halotools/mock_observables/two_point_clustering/tpcf_estimators.py:12-183).

The pair-count arrays may be NumPy arrays or CUDA tensors (the caller layer
keeps its counts on the device): the same expressions, in the reference's
operand order, run on either.  Elementwise FP64 +, -, *, / are correctly
rounded IEEE operations in both libraries, so the device results equal the
NumPy ones bit for bit.  Counts arrive as float64 (integer counts below 2**53
convert exactly, and the reference's int64 true divisions convert the same way).
"""
import numpy as np
import torch

from ...custom_exceptions import HalotoolsError

__all__ = ("_TP_estimator", "_TP_estimator_crossx", "_list_estimators", "_TP_estimator_requirements",
           "_test_for_zero_division")


def _coef(c, like):
    """A per-sample factor (NumPy, length 1 or Njack) in the container of `like`."""
    if isinstance(like, torch.Tensor):
        if c.size == 1:
            return float(c[0])  # a Python float is an FP64 scalar operand for a float64 tensor
        return torch.as_tensor(c, dtype=torch.float64, device=like.device)
    return c


def _mult_factory(ND1, ND2, NR1, NR2):
    Ns = np.array([len(ND1), len(ND2), len(NR1), len(NR2)])
    if np.any(Ns > 1):
        # jackknife samples: scale each row by its own factor (tpcf_estimators.py:26-33)
        def mult(x, y):
            x = _coef(x, y)
            return (x * y.T).T
    else:
        def mult(x, y):
            return _coef(x, y) * y
    return mult


def _first_or_all(xi):
    return xi[0] if tuple(xi.shape)[0] == 1 else xi


def _TP_estimator(DD, DR, RR, ND1, ND2, NR1, NR2, estimator):
    """tpcf_estimators.py:12-73."""
    ND1 = np.atleast_1d(ND1)
    ND2 = np.atleast_1d(ND2)
    NR1 = np.atleast_1d(NR1)
    NR2 = np.atleast_1d(NR2)
    mult = _mult_factory(ND1, ND2, NR1, NR2)
    _test_for_zero_division(DD, DR, RR, ND1, ND2, NR1, NR2, estimator)
    if estimator == "Natural":
        factor = ND1 * ND2 / (NR1 * NR2)
        xi = mult(1.0 / factor, DD / RR) - 1.0  # DD/RR - 1
    elif estimator == "Davis-Peebles":
        factor = ND1 * ND2 / (ND1 * NR2)
        xi = mult(1.0 / factor, DD / DR) - 1.0  # DD/DR - 1
    elif estimator == "Hewett":
        factor1 = ND1 * ND2 / (NR1 * NR2)
        factor2 = ND1 * NR2 / (NR1 * NR2)
        xi = mult(1.0 / factor1, DD / RR) - mult(1.0 / factor2, DR / RR)  # (DD - DR)/RR
    elif estimator == "Hamilton":
        xi = (DD * RR) / (DR * DR) - 1.0  # DD RR / DR^2 - 1
    elif estimator == "Landy-Szalay":
        factor1 = ND1 * ND2 / (NR1 * NR2)
        factor2 = ND1 * NR2 / (NR1 * NR2)
        xi = mult(1.0 / factor1, DD / RR) - mult(1.0 / factor2, 2.0 * DR / RR) + 1.0  # (DD - 2DR + RR)/RR
    else:
        raise ValueError("unsupported estimator!")
    return _first_or_all(xi)


def _TP_estimator_crossx(DD, D1R, D2R, RR, ND1, ND2, NR1, NR2, estimator):
    """tpcf_estimators.py:76-129."""
    ND1 = np.atleast_1d(ND1)
    ND2 = np.atleast_1d(ND2)
    NR1 = np.atleast_1d(NR1)
    NR2 = np.atleast_1d(NR2)
    mult = _mult_factory(ND1, ND2, NR1, NR2)
    _test_for_zero_division(DD, D1R, RR, ND1, ND2, NR1, NR2, estimator)
    _test_for_zero_division(DD, D2R, RR, ND1, ND2, NR1, NR2, estimator)
    if estimator == "Natural":
        factor = ND1 * ND2 / (NR1 * NR2)
        xi = mult(1.0 / factor, DD / RR) - 1.0
    elif estimator == "Hamilton":
        xi = (DD * RR) / (D1R * D2R) - 1.0
    elif estimator == "Landy-Szalay":
        factor1 = ND1 * ND2 / (NR1 * NR2)
        factor2 = ND1 * NR2 / (NR1 * NR2)
        term1 = mult(1.0 / factor1, DD / RR)
        term2 = mult(1.0 / factor2, D1R / RR)
        term3 = mult(1.0 / factor2, D2R / RR)
        xi = term1 - term2 - term3 + 1.0
    else:
        msg = "{0} estimator is not supported for cross-correlations"
        raise ValueError(msg.format(estimator))
    return _first_or_all(xi)


def _list_estimators():
    """tpcf_estimators.py:132-137."""
    return ["Natural", "Davis-Peebles", "Hewett", "Hamilton", "Landy-Szalay"]


def _TP_estimator_requirements(estimator):
    """tpcf_estimators.py:140-169: (do_DD, do_DR, do_RR)."""
    table = {"Natural": (True, False, True), "Davis-Peebles": (True, True, False), "Hewett": (True, True, True),
             "Hamilton": (True, True, True), "Landy-Szalay": (True, True, True)}
    if estimator not in table:
        msg = "Input `estimator` must be one of the following:{0}".format(_list_estimators())
        raise HalotoolsError(msg)
    return table[estimator]


def _any_zero(a):
    if a is None:
        return False
    if isinstance(a, torch.Tensor):
        return bool((a == 0).any())
    return bool(np.any(a == 0))


def _test_for_zero_division(DD, DR, RR, ND1, ND2, NR1, NR2, estimator):
    """tpcf_estimators.py:172-183."""
    zero_msg = ("When calculating the two-point function, there was at least one \n"
                "separation bin with zero {0} pairs. Since the ``{1}`` estimator you chose \n"
                "divides by {0}, you will have at least one NaN returned value.\n"
                "Most likely, the innermost separation bin is the problem.\n"
                "Try increasing the number of randoms and/or using broader bins.\n"
                "To estimate the number of required randoms, the following expression \n"
                "for the expected number of pairs inside a sphere of radius ``r`` may be useful:\n\n"
                "<Npairs> = (Nran_tot)*(4pi/3)*(r/Lbox)^3 \n\n")
    estimators_dividing_by_rr = ("Natural", "Davis-Peebles", "Hewett", "Landy-Szalay")
    if (estimator in estimators_dividing_by_rr) and _any_zero(RR):
        raise ValueError(zero_msg.format("RR", estimator))
    estimators_dividing_by_dr = ("Hamilton",)
    if (estimator in estimators_dividing_by_dr) and _any_zero(DR):
        raise ValueError(zero_msg.format("DR", estimator))
