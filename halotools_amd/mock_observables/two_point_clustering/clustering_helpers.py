"""Argument helpers of the two-point functions
(halotools/mock_observables/two_point_clustering/clustering_helpers.py:16-95)."""
from warnings import warn

import numpy as np
import torch

from ..mock_observables_helpers import enforce_sample_has_correct_shape

__all__ = ("verify_tpcf_estimator", "process_optional_input_sample2", "tpcf_estimator_dd_dr_rr_requirements")

# estimator name -> (do_DD, do_DR, do_RR), clustering_helpers.py:16-23
tpcf_estimator_dd_dr_rr_requirements = {
    "Natural": (True, False, True),
    "Davis-Peebles": (True, True, False),
    "Hewett": (True, True, True),
    "Hamilton": (True, True, True),
    "Landy-Szalay": (True, True, True),
}


def verify_tpcf_estimator(estimator):
    """clustering_helpers.py:26-41."""
    available_estimators = list(tpcf_estimator_dd_dr_rr_requirements.keys())
    if estimator in available_estimators:
        return estimator
    msg = (u"Your estimator ``{0}`` \n"
           "is not in the list of available estimators:\n {1}".format(estimator, available_estimators))
    raise ValueError(msg)


def _all_equal(a, b):
    if isinstance(a, torch.Tensor) or isinstance(b, torch.Tensor):
        if not (isinstance(a, torch.Tensor) and isinstance(b, torch.Tensor)) or a.device != b.device:
            a = a.cpu().numpy() if isinstance(a, torch.Tensor) else a
            b = b.cpu().numpy() if isinstance(b, torch.Tensor) else b
            return bool(np.all(a == b))
        return bool((a == b).all())
    return bool(np.all(a == b))


def process_optional_input_sample2(sample1, sample2, do_cross, ndim=3):
    """clustering_helpers.py:44-95: sample2 None or equal to sample1 -> auto only."""
    if sample2 is None:
        sample2 = sample1
        _sample1_is_sample2 = True
    else:
        sample2 = enforce_sample_has_correct_shape(sample2, ndim=ndim)
        if tuple(sample1.shape) != tuple(sample2.shape):
            _sample1_is_sample2 = False
        else:
            if _all_equal(sample1, sample2):
                _sample1_is_sample2 = True
                if do_cross:
                    msg = (u"\n `sample1` and `sample2` are exactly the same, \n"
                           "only the auto-correlation will be returned.\n")
                    warn(msg)
                    do_cross = False
            else:
                _sample1_is_sample2 = False
    return sample2, _sample1_is_sample2, do_cross
