"""pes's observables as product code (mpci_obs_weights, host side of the C
ABI; no GPU needed): per-state weight rows for the ground population,
yield, excited population and norm of src/pes/pes.cpp:48-262, checked
against the oracle (or_observables, or_observables_1e) and against the
hand-worked cases that pin the oracle."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def mpci():
    import mpci as m
    return m


def _state(N, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(2 * N)
    return x / np.linalg.norm(x)


def _obs(W, x):
    return W @ (x[0::2] ** 2 + x[1::2] ** 2)


def test_weights_2e_match_oracle(mpci, oracle):
    from mpci import synth
    s = [40, 33, 27, 35]
    E, D = synth.system_arrays(s, seed=21)
    o = oracle.System(E, D)
    W = mpci.obs_weights(E, 2, -2.0)
    assert W.shape == (mpci.OBS_ROWS, sum(s))
    for seed in range(3):
        x = _state(sum(s), seed)
        got = _obs(W, x)
        ref = o.observables(x, -2.0)
        np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-16)
    # the yield depends on the threshold through the energies only
    W2 = mpci.obs_weights(E, 2, 0.5)
    assert W2[mpci.OBS_YIELD].sum() < W[mpci.OBS_YIELD].sum()


def test_weights_1e_match_oracle(mpci, oracle):
    from mpci import synth
    s = [30, 45, 40]  # s_0 <= s_1: the bins read l = 1 energies to s_0 - 1
    E, D = synth.system_arrays(s, seed=22, e_lo=-0.5, e_hi=2.0)
    o = oracle.System(E, D)
    W = mpci.obs_weights(E, 1)
    assert W[mpci.OBS_YIELD].sum() > 0
    for seed in range(3):
        x = _state(sum(s), seed)
        got = _obs(W, x)
        ref = o.observables_1e(x)
        np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-16)


def test_weights_hand_cases(mpci):
    # the 2e hand case of the oracle's pin (two blocks, threshold -2.0)
    E = [np.array([-2.9, -2.1, -1.5, 0.3]), np.array([-2.5, -1.9, 2.0])]
    W = mpci.obs_weights(E, 2, -2.0)
    assert list(np.nonzero(W[mpci.OBS_YIELD])[0]) == [2, 5]
    assert list(np.nonzero(W[mpci.OBS_EXCITED])[0]) == [1, 2, 5]
    assert list(np.nonzero(W[mpci.OBS_GROUND])[0]) == [0]
    assert np.all(W[mpci.OBS_NORM] == 1.0)
    # the 1e binning case of the oracle's pin
    E = [np.array([-2.0, -0.25, 0.0, 0.4375, 1.0]),
         np.array([-0.5, -0.125, 0.125, 0.25, 0.75, 1.0]),
         np.array([0.0625, 0.15625, 0.1875])]
    W = mpci.obs_weights(E, 1)
    assert list(np.nonzero(W[mpci.OBS_YIELD])[0]) == [3, 7, 8, 12]
    assert not W[mpci.OBS_EXCITED].any()


def test_weights_refuse_bad_input(mpci):
    E = [np.array([-1.0, 0.5, 1.0]), np.array([0.1, 0.2])]
    with pytest.raises(mpci.MpciError):
        mpci.obs_weights(E, 1)  # s_0 > s_1: pes would read past l = 1
    with pytest.raises(mpci.MpciError):
        mpci.obs_weights(E[:1], 1)  # no l = 1 block to bin onto
    with pytest.raises(mpci.MpciError):
        mpci.obs_weights(E, 3)
