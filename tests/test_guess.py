"""Learned SCC start charges: the numpy oracle of the descriptor and network
(CPU), the HIP descriptor/forward/backward kernels against it (GPU), and what a
trained network gains over the standard start (GPU)."""
import numpy as np
import pytest

import guess_oracle as go
from gfn1_amd import geometry

ANG = geometry.ANGSTROM_TO_BOHR


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


# ---------------------------------------------------------------- CPU: oracle

def test_oracle_backward_matches_finite_differences():
    rng = np.random.default_rng(5)
    theta = rng.normal(size=go.NPARAM) * 0.2
    norm = np.concatenate([rng.normal(size=go.NFEAT) * 0.1, 1.0 + rng.random(go.NFEAT)])
    feat = rng.random((7, go.NFEAT))
    labels = rng.normal(size=(7, 2)) * 0.3
    loss, grad = go.loss_and_gradient(theta, norm, feat, labels)
    # one index in every parameter block and a random sample
    picks = [0, go.OFF_B1 + 3, go.OFF_W2 + 77, go.OFF_B2 + 5, go.OFF_W3 + 70, go.OFF_B3 + 1]
    picks += list(rng.integers(0, go.NPARAM, 20))
    h = 1e-6
    for p in picks:
        tp, tm = theta.copy(), theta.copy()
        tp[p] += h
        tm[p] -= h
        fd = (go.loss_and_gradient(tp, norm, feat, labels)[0] - go.loss_and_gradient(tm, norm, feat, labels)[0]) / (2 * h)
        assert abs(fd - grad[p]) <= 1e-7 + 1e-6 * abs(fd), p
    # the loss is the mean squared error over both outputs of every atom
    y = go.forward(theta, norm, feat)
    assert loss == pytest.approx(np.mean((y - labels) ** 2), rel=1e-12)


def test_descriptor_invariances_and_cutoff():
    z, x = geometry.noisy_snapshot(3, kmax=8)
    f0 = go.features(z, x)
    rng = np.random.default_rng(1)
    rot = random_rotation(rng)
    x2 = x @ rot.T + np.array([3.0, -1.0, 7.0])
    assert np.allclose(go.features(z, x2), f0, atol=1e-12)
    perm = rng.permutation(len(z))
    assert np.allclose(go.features(z[perm], x[perm]), f0[perm], atol=1e-12)
    # an atom beyond the cutoff of everything changes nothing for the others
    far = np.vstack([x, x.mean(0) + np.array([0.0, 0.0, 40.0 * ANG])])
    f_far = go.features(np.append(z, 8), far)
    assert np.allclose(f_far[:-1], f0, atol=1e-15)
    assert np.allclose(f_far[-1, 4:], 0.0)


def test_descriptor_of_an_isolated_pair():
    r = 1.3  # Angstrom
    f = go.features([6, 1], [[0.0, 0.0, 0.0], [0.0, 0.0, r * ANG]])
    mu = 0.5 + 0.3 * np.arange(16)
    fc = 0.5 * (np.cos(np.pi * r / 5.0) + 1.0)
    g = np.exp(-(r - mu) ** 2 / (2 * 0.3 ** 2)) * fc
    assert np.array_equal(f[0, :4], [0, 1, 0, 0]) and np.array_equal(f[1, :4], [1, 0, 0, 0])
    assert np.allclose(f[0, 4:20], g, rtol=1e-14)         # C sees one H
    assert np.allclose(f[1, 20:36], g, rtol=1e-14)        # H sees one C
    assert np.allclose(np.delete(f[0, 4:], np.s_[0:16]), 0.0)
    assert np.allclose(np.delete(f[1, 4:], np.s_[16:32]), 0.0)


# ---------------------------------------------------------------- GPU: kernels

def _systems(count, first, kmax=10):
    return geometry.noisy_snapshots(count, first=first, kmax=kmax)


@pytest.mark.gpu
def test_guess_features_match_oracle():
    from gfn1_amd import guess_features
    systems = _systems(3, 40) + [(np.array([6, 1, 1, 1, 35]), np.array(
        [[0, 0, 0], [1.0, 1.0, 1.0], [-1.0, -1.0, 1.0], [1.0, -1.0, -1.0], [-2.1, 2.1, -2.1]]) * ANG)]
    print("case guess features", flush=True)
    f = guess_features(systems)
    ref = np.vstack([go.features(z, x) for z, x in systems])
    assert f.shape == ref.shape
    assert np.allclose(f, ref, rtol=2e-6, atol=2e-6)


@pytest.mark.gpu
def test_guess_loss_gradient_and_prediction_match_oracle():
    from gfn1_amd import GuessModel
    systems = _systems(5, 60)
    rng = np.random.default_rng(2)
    labels = [rng.normal(size=2 * len(z)) * 0.2 for z, _ in systems]
    m = GuessModel(seed=11)
    theta, norm = m.params()
    print("case guess loss/gradient", flush=True)
    loss, grad = m.loss_gradient(systems, labels)
    feat = np.vstack([go.features(z, x) for z, x in systems])
    lab = np.concatenate(labels).reshape(-1, 2)
    ref_loss, ref_grad = go.loss_and_gradient(theta, norm, feat, lab)
    assert loss == pytest.approx(ref_loss, rel=1e-5)
    assert np.linalg.norm(grad - ref_grad) <= 1e-4 * np.linalg.norm(ref_grad)
    y = go.forward(theta, norm, feat)
    q = m.predict(systems)
    off = 0
    for (z, _), qs in zip(systems, q):
        ref = go.start_vector(y[off:off + len(z)], 0.0)
        off += len(z)
        assert abs(qs.sum()) < 1e-9
        assert np.allclose(qs, ref, atol=5e-6)


@pytest.mark.gpu
def test_trained_guess_cuts_iterations_and_keeps_the_result():
    from gfn1_amd import AnalyticGradientOptions, BatchEngine, ElectronicOptions, GuessModel, basis_counts
    tight = AnalyticGradientOptions(electronic=ElectronicOptions(energy_tolerance=1e-10, charge_tolerance=1e-7))
    train = _systems(600, 0, kmax=12)
    print("case guess training: labels", flush=True)
    eng = BatchEngine(train, options=tight).run()
    assert all(eng.status(i) == 0 for i in range(len(train)))
    labels = [eng.shell_charges(i, basis_counts(z)[0]) for i, (z, _) in enumerate(train)]
    m = GuessModel(seed=3)
    print("case guess training: fit", flush=True)
    loss = m.train(train, labels, epochs=30, batch_atoms=1024, learning_rate=2e-3)
    assert loss[-1] < 0.2 * loss[0]
    test = _systems(64, 10_000, kmax=12)
    print("case guess training: evaluation", flush=True)
    plain = BatchEngine(test, options=tight).run()
    learned = BatchEngine(test, options=tight, guess=m).run()
    it_plain = np.array([plain.iterations(i) for i in range(len(test))])
    it_learned = np.array([learned.iterations(i) for i in range(len(test))])
    for i in range(len(test)):
        assert plain.status(i) == 0 and learned.status(i) == 0
        assert abs(plain.energy(i) - learned.energy(i)) <= 1e-8
        assert np.abs(plain.gradient(i) - learned.gradient(i)).max() <= 1e-7
    print("iterations: standard start %.2f, learned start %.2f" % (it_plain.mean(), it_learned.mean()), flush=True)
    assert it_learned.mean() < it_plain.mean() - 1.0
    # the batch keeps working without the model
    learned.set_guess(None)
    learned.run()
    assert np.array_equal([learned.iterations(i) for i in range(len(test))], it_plain)
