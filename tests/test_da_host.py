"""CPU checks of the assimilation's host arithmetic against closed forms:
the logit map of the hyper-parameters and its inverse
(data_assimilation/DataAssimilation.py:110-140), the initial logit grid
(:204-252), the surrogate's window-base BOLD against the reference BOLD
model's own output (models/bold_model_pytorch.py:76-77, golden vectors) and
the surrogate NRMSE, ||sur - sim||_2 / (sqrt(n) range(sim))."""
import os

import numpy as np
import pytest
import torch

from digital_twin_brain_amd.data_assimilation import DataAssimilation
from digital_twin_brain_amd.surrogate_da import nrmse, window_base


def _da(E, populations, voxels):
    """A DataAssimilation without a network: only the ensemble bookkeeping the
    hyper-parameter methods read."""
    da = DataAssimilation.__new__(DataAssimilation)
    da._ensemble_number = E
    da._da_populations = populations
    da._da_voxels = voxels
    return da


def test_logit_map_and_its_inverse():
    lo = torch.tensor([0.5, 0.1, 1.0, 2.0], dtype=torch.float64)
    hi = torch.tensor([2.0, 0.3, 5.0, 2.5], dtype=torch.float64)
    x = lo + (hi - lo) * torch.rand(1000, 4, dtype=torch.float64, generator=torch.Generator().manual_seed(1))
    y = DataAssimilation.log_torch(x, lo, hi)
    # 10 log((x - lo) / (hi - x)), the logit of (x - lo) / (hi - lo) scaled by 10
    u = (x - lo) / (hi - lo)
    assert torch.allclose(y, 10 * torch.log(u / (1 - u)), rtol=1e-12, atol=1e-12)
    assert torch.allclose(DataAssimilation.sigmoid_torch(y, lo, hi), x, rtol=1e-12, atol=1e-12)
    assert torch.allclose(DataAssimilation.sigmoid_torch(torch.zeros(3, 4, dtype=torch.float64), lo, hi),
                          ((lo + hi) / 2).expand(3, 4))


@pytest.mark.parametrize("E,P,H", [(64, 4, 2), (256, 90, 4), (7, 3, 1)])
def test_initial_grid_is_the_middle_third_evenly_spaced(E, P, H):
    """The logit grid 10 log((1 + t) / (1 - t)), t evenly spaced over [-1/3, 1/3],
    puts each member at lo + (hi - lo) (1 + t) / 2: the middle third of the
    prior range, evenly spaced; each (population, parameter) column is a
    permutation of that grid."""
    da = _da(E, E * P, E * P)
    rng = np.random.default_rng(3)
    lo = torch.from_numpy(rng.uniform(0.1, 1.0, size=(P, H))).float()
    hi = lo + torch.from_numpy(rng.uniform(0.5, 2.0, size=(P, H))).float()
    hp = da.hp_random_initialize(lo, hi, rng=np.random.default_rng(5)).reshape(E, P, H).double()
    t = np.linspace(-1 / 3, 1 / 3, E)
    for p in range(P):
        for h in range(H):
            want = lo[p, h].double() + (hi[p, h] - lo[p, h]).double() * (1 + torch.from_numpy(t)) / 2
            got = torch.sort(hp[:, p, h]).values
            assert torch.allclose(got, want, rtol=1e-5, atol=1e-6), (p, h)
    if E > 8 and H > 1:
        # members are shuffled independently per parameter column
        assert not torch.equal(torch.argsort(hp[:, 0, 0]), torch.argsort(hp[:, 0, 1]))


def test_nrmse_closed_forms():
    rng = np.random.default_rng(7)
    sim = rng.normal(size=(16, 90))
    r = sim.max() - sim.min()
    assert nrmse(sim, sim) == 0.0
    assert nrmse(sim + 0.03, sim) == pytest.approx(0.03 / r, rel=1e-12)
    d = rng.normal(size=sim.shape)
    assert nrmse(sim + d, sim) == pytest.approx(np.sqrt((d ** 2).mean()) / r, rel=1e-12)
    assert nrmse(sim + d, sim, scale=2.0) == pytest.approx(np.sqrt((d ** 2).mean()) / 2.0, rel=1e-12)
    # the range is the simulation's, not the surrogate's
    assert nrmse(10 * sim, sim) == pytest.approx(9 * np.sqrt((sim ** 2).mean()) / r, rel=1e-12)


def test_window_base_bold_equals_the_reference_output_of_the_same_state():
    """With nothing changing, a window ends at its start state: window_base's
    BOLD of a state (q, v) must be what the reference BOLD model output for
    that state.  tests/golden/bold_v18.npz holds the reference's state every
    10th step (state[k], after step 10 k + 10) beside its output (out)."""
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "bold_v18.npz"))
    assert float(z["E_0"]) == 0.8 and float(z["V_0"]) == 0.02
    st = z["state"]  # [K, 4 (s, q, v, f_in), V]
    K, _, V = st.shape
    X = np.zeros((K, V, 9))
    X[..., 4:8] = st.transpose(0, 2, 1)
    X[..., 8] = np.arange(K * V).reshape(K, V)  # the previous act, carried through
    base = window_base(torch.from_numpy(X.reshape(K, -1)), V).reshape(K, V, 6).numpy()
    np.testing.assert_allclose(base[..., 5], z["out"][9::10][:K], rtol=1e-12, atol=1e-15)
    np.testing.assert_array_equal(base[..., 1:5], X[..., 4:8])
    np.testing.assert_array_equal(base[..., 0], X[..., 8])
    no_act = window_base(torch.from_numpy(X.reshape(K, -1)), V, use_act=False).reshape(K, V, 6).numpy()
    assert np.all(no_act[..., 0] == 0)
