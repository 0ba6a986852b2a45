"""Drop-in controller ``Model`` (``multitask/nn.py:28-196``).

Two-layer MLP ``act(W2 act(W1 in + b1) + b2)`` evaluated per step by
``dl_mlp_forward`` (HIP); its adjoint ``dl_mlp_backward`` is recorded on the
active ``Tape``.  Optimiser state and updates follow ``nn.py:163-196`` (Adam
with bias correction by ``iter + 1``; the reference's betas are 0.9/0.9).
"""

import math
import pickle

import numpy as np
import torch

from . import _native as nat
from .fields import Field, record


class _ArrayUnpickler(pickle.Unpickler):
    """Unpickler for lists of numpy arrays: only the globals numpy's array
    pickles use (numpy 1.x and 2.x module paths) resolve."""

    _ALLOWED = {("numpy.core.multiarray", "_reconstruct"), ("numpy._core.multiarray", "_reconstruct"),
                ("numpy.core.multiarray", "scalar"), ("numpy._core.multiarray", "scalar"),
                ("numpy", "ndarray"), ("numpy", "dtype")}

    def find_class(self, module, name):
        if (module, name) in self._ALLOWED:
            return super().find_class(module, name)
        raise pickle.UnpicklingError(f"weights file refers to {module}.{name}: refused")


class Model:
    def __init__(self, config, steps, batch_size, n_input, n_output, input, output,
                 n_hidden=64, method="adam"):
        cfg = config.get_config()
        self.adam_b1 = cfg["nn"]["adam_b1"]
        self.adam_b2 = cfg["nn"]["adam_b2"]
        self.learning_rate = cfg["nn"]["learning_rate"]
        self.dim = cfg["robot"]["dim"]
        self.activation = cfg["nn"].get("activation", "sin")
        self.init = cfg["nn"].get("init", "uniform")
        self.steps = steps
        self.batch_size = batch_size
        self.n_input = n_input
        self.n_output = n_output
        self.n_hidden = n_hidden
        self.method = method
        self.input = input
        self.output_act = output
        dt, dev = input.data.dtype, input.data.device
        mk = lambda shape, name: Field(shape, dt, dev, name=name)
        self.hidden_act = mk((steps + 1, batch_size, n_hidden), "hidden_act")
        self.weights1 = mk((n_hidden, n_input), "W1")
        self.bias1 = mk((n_hidden,), "b1")
        self.weights2 = mk((n_output, n_hidden), "W2")
        self.bias2 = mk((n_output,), "b2")
        self.weights = [self.weights1, self.weights2, self.bias1, self.bias2]
        self.adam_weights = None
        if method == "adam":
            self.adam_weights = [(torch.zeros_like(w.data), torch.zeros_like(w.data))
                                 for w in self.weights]
        self.TNS = 0.0
        self._work = torch.zeros(batch_size * (2 * n_hidden + n_output), dtype=dt, device=dev)

    @property
    def _act(self):
        return nat.ACT_SIN if self.activation == "sin" else nat.ACT_TANH

    def weights_init(self, seed=None):
        """nn.py:59-67: U(-sqrt(6/fan_in), +sqrt(6/fan_in)) weights, biases
        untouched; with init="normal", N(0, 1/fan_in) and zero biases."""
        g = torch.Generator(device="cpu")
        g.manual_seed(0 if seed is None else int(seed))
        for w, fan in ((self.weights1, self.n_input), (self.weights2, self.n_hidden)):
            if self.init == "normal":
                r = torch.randn(w.shape, generator=g, dtype=torch.float64) / math.sqrt(fan)
            else:
                r = (torch.rand(w.shape, generator=g, dtype=torch.float64) * 2 - 1) * math.sqrt(6 / fan)
            w.data.copy_(r)
        if self.init == "normal":
            self.bias1.fill(0)
            self.bias2.fill(0)

    def set_weights(self, W1, b1, W2, b2):
        for f, v in zip((self.weights1, self.bias1, self.weights2, self.bias2), (W1, b1, W2, b2)):
            f.data.copy_(torch.as_tensor(v))

    def clear_adam(self):
        if self.adam_weights is not None:
            for m, v in self.adam_weights:
                m.zero_()
                v.zero_()

    def clear(self):
        self.hidden_act.fill(0)

    def forward(self, t):
        """nn1(t) + nn2(t) (nn.py:130-150)."""
        B, I, H, O = self.batch_size, self.n_input, self.n_hidden, self.n_output
        dc = nat.dtype_code(self.input.data.dtype)
        W = [w.data for w in (self.weights1, self.bias1, self.weights2, self.bias2)]
        nat.call("dl_mlp_forward", dc, self._act, B, I, H, O, nat.ptr(self.input.at(t)),
                 *map(nat.ptr, W), nat.ptr(self.hidden_act.at(t)),
                 nat.ptr(self.output_act.at(t)), nat.stream())

        def adj():
            G = [w.grad.data for w in (self.weights1, self.bias1, self.weights2, self.bias2)]
            nat.call("dl_mlp_backward", dc, self._act, B, I, H, O, nat.ptr(self.input.at(t)),
                     *map(nat.ptr, W), nat.ptr(self.output_act.grad.at(t)),
                     nat.ptr(self.input.grad.at(t)), *map(nat.ptr, G), nat.ptr(self._work),
                     nat.stream())
        record(adj)

    def dump_weights(self, name="save.pkl"):
        """Weights as the list [W1, W2, b1, b2] of numpy arrays (nn.py:152-156)."""
        with open(name, "wb") as fh:
            pickle.dump([w.to_numpy() for w in self.weights], fh)

    def load_weights(self, name="save.pkl"):
        """Load a ``[W1, W2, b1, b2]`` weights pickle (nn.py:158-161), written
        by ``dump_weights`` or by the reference.  The file is read with a
        restricted unpickler that resolves only numpy's array-reconstruction
        globals, so a file naming any other callable is refused instead of run."""
        with open(name, "rb") as fh:
            vals = _ArrayUnpickler(fh).load()
        if not isinstance(vals, (list, tuple)) or len(vals) != len(self.weights):
            raise ValueError(f"{name}: expected a list of {len(self.weights)} arrays")
        for w, val in zip(self.weights, vals):
            if not isinstance(val, np.ndarray) or tuple(val.shape) != tuple(w.view.shape):
                raise ValueError(f"{name}: array of shape {getattr(val, 'shape', None)} "
                                 f"for a field of shape {tuple(w.view.shape)}")
            w.from_numpy(val)

    def gradient_update(self, iter=0):
        """nn.py:163-171, 184-196."""
        with torch.no_grad():
            if self.method == "adam":
                b1, b2, lr = self.adam_b1, self.adam_b2, self.learning_rate
                for w, (m, v) in zip(self.weights, self.adam_weights):
                    g = w.grad.data
                    m.mul_(b1).add_((1 - b1) * g)
                    v.mul_(b2).add_((1 - b2) * g * g)
                    m_cap = m / (1 - b1 ** (iter + 1))
                    v_cap = v / (1 - b2 ** (iter + 1))
                    w.data.sub_((lr * m_cap) / (torch.sqrt(v_cap) + 1e-8))
            else:
                for w in self.weights:
                    w.data.sub_(w.grad.data * self.learning_rate)

    def get_TNS(self):
        """Total squared norm of the weight gradients (nn.py:173-182)."""
        self.TNS = float(sum((w.grad.data.double() ** 2).sum().item() for w in self.weights))
        return self.TNS
