"""The no-allocation learned-start prediction (gfn1_guess_predict_noalloc)
against gfn1_guess_predict through the C ABI: identical start vectors on two
calls in a row, the second with a smaller input, and the size limit."""
import os
import sys

import numpy as np
import pytest

from gfn1_amd import GuessModel, NativeError, geometry

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_noalloc_prediction_matches_the_allocating_one():
    model = GuessModel.load(os.path.join(ROOT, "gfn1-rs_amd", "data", "guess_mlp.bin"))
    big = geometry.mixed_batch(64, first_seed=1000)
    # an element the model does not cover (P): that system keeps the zero start
    z_ph = np.array([15, 1, 1, 1])
    x_ph = np.array([[0, 0, 0], [1.193, 0, 0.768], [-0.5965, 1.0333, 0.768], [-0.5965, -1.0333, 0.768]]) * 1.8897261246257702
    small = geometry.mixed_batch(5, first_seed=2000, kmax=6) + [(z_ph, x_ph)]
    note("case no-alloc prediction: %d systems, %d atoms" % (len(big), sum(len(z) for z, _ in big)))
    got = model.predict(big, noalloc=True)
    want = model.predict(big)
    assert all(np.array_equal(g, w) for g, w in zip(got, want))
    note("case no-alloc prediction, smaller second call: %d systems" % len(small))
    got = model.predict(small, noalloc=True)
    want = model.predict(small)
    assert all(np.array_equal(g, w) for g, w in zip(got, want))
    assert np.all(got[-1] == 0.0) and np.abs(got[0]).max() > 0.0
    # a third call, back to the large input, still agrees (the scratch keeps no state)
    got = model.predict(big[:7], noalloc=True)
    assert all(np.array_equal(g, w) for g, w in zip(got, model.predict(big[:7])))


def test_noalloc_prediction_rejects_inputs_beyond_its_limit():
    model = GuessModel.load(os.path.join(ROOT, "gfn1-rs_amd", "data", "guess_mlp.bin"))
    n = GuessModel.NOALLOC_MAX_ATOMS + 1
    x = np.zeros((n, 3))
    x[:, 0] = np.arange(n) * 10.0
    with pytest.raises(NativeError) as exc:
        model.predict([(np.ones(n, dtype=np.int32), x)], noalloc=True)
    assert "at most" in str(exc.value)
