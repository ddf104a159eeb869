"""The reference's closed-form known answers (tests/golden/known_answers.json,
written by tests/golden/make_known_answers.py with a citation per case) against
the oracle (CPU) and the HIP path (GPU)."""
import json
import os

import numpy as np
import pytest

from oracle import mesh as ORC
from tests.helpers import generate_3d_regular_mesh, generate_locus_of_3d_points

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = json.load(open(os.path.join(HERE, "golden", "known_answers.json")))["cases"]


def _sample(spec):
    if spec["kind"] == "grid":
        return generate_3d_regular_mesh(spec["npts_per_dim"])
    return generate_locus_of_3d_points(spec["npts"], xc=spec["xc"], yc=spec["yc"], zc=spec["zc"], seed=spec["seed"])


def _call(module, case):
    s1, s2 = _sample(case["sample1"]), _sample(case["sample2"])
    args = [np.array(v) if isinstance(v, list) else v for v in case["args"].values()]
    kw = dict(case["kwargs"])
    if module is ORC and "num_threads" not in kw:
        kw["num_threads"] = 2
    return np.asarray(getattr(module, case["function"])(s1, s2, *args, **kw))


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_oracle_known_answers(case):
    assert np.array_equal(_call(ORC, case), np.array(case["expected"])), case["reference"]


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_gpu_known_answers(gpu, case):
    import halotools_amd.mock_observables as mo
    assert np.array_equal(_call(mo, case), np.array(case["expected"])), case["reference"]


def test_fixture_is_current():
    # the committed fixture is what the generator writes
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_known_answers", os.path.join(HERE, "golden",
                                                                                      "make_known_answers.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    assert json.loads(json.dumps({"cases": mod.CASES})) == {"cases": CASES}
