"""Model weight files: the reference's ``[W1, W2, b1, b2]`` pickle
(nn.py:152-161) round trip, and refusal of a pickle that names anything but
numpy's array globals (no code from the file runs)."""

import pickle

import numpy as np
import pytest
import torch

import diffloco as dl
from diffloco.fields import Field
from diffloco.nn import Model


def _model(seed=None):
    cfg = dl.quadruped_config(batch_size=2, steps=10, window=5, precision="f64")
    c = cfg.get_config()
    I, O = c["nn"]["n_input_states"], c["robot"]["n_actuators"]
    inp = Field((11, 2, I), torch.float64, "cpu", name="input")
    out = Field((11, 2, O), torch.float64, "cpu", name="actuation")
    m = Model(cfg, 10, 2, I, O, inp, out, n_hidden=64)
    if seed is not None:
        m.weights_init(seed=seed)
    return m


def test_pickle_round_trip_reference_format(tmp_path):
    a = _model(seed=3)
    a.bias1.data.copy_(torch.linspace(-1, 1, 64, dtype=torch.float64))
    f = tmp_path / "save.pkl"
    a.dump_weights(str(f))
    with open(f, "rb") as fh:
        vals = pickle.load(fh)          # our own file: the format is a plain list of 4 arrays
    assert [v.shape for v in vals] == [(64, 179), (4, 64), (64,), (4,)]
    b = _model()
    b.load_weights(str(f))
    for x, y in zip(a.weights, b.weights):
        assert torch.equal(x.data, y.data)


class _Trap:
    def __init__(self, path):
        self.path = path

    def __reduce__(self):
        return (open, (self.path, "w"))


def test_pickle_naming_other_callables_is_refused(tmp_path):
    flag = tmp_path / "ran"
    f = tmp_path / "evil.pkl"
    with open(f, "wb") as fh:
        pickle.dump([np.zeros((64, 179)), _Trap(str(flag)), np.zeros(64), np.zeros(4)], fh)
    with pytest.raises(pickle.UnpicklingError):
        _model().load_weights(str(f))
    assert not flag.exists()


def test_wrong_shapes_are_refused(tmp_path):
    f = tmp_path / "bad.pkl"
    with open(f, "wb") as fh:
        pickle.dump([np.zeros((64, 10)), np.zeros((4, 64)), np.zeros(64), np.zeros(4)], fh)
    with pytest.raises(ValueError):
        _model().load_weights(str(f))
