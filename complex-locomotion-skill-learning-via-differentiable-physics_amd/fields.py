"""Taichi-style fields on the GPU and a Tape that replays hand-written adjoints.

The reference keeps every quantity in Taichi fields indexed ``[t, k, i]``
(``multitask/solver_mass_spring.py:25-56``) and differentiates with
``ti.ad.Tape(loss)`` (``multitask/diffphy_trainer.py:246``), which zeroes all
gradients on entry, and on exit seeds ``loss.grad = 1`` and runs the adjoint
of every recorded kernel in reverse order.  ``Field`` and ``Tape`` provide the
same surface over torch device tensors; each forward call records the closure
that launches its adjoint kernel.
"""

import weakref

import numpy as np
import torch

_tapes = []
_grad_fields = weakref.WeakSet()


class Field:
    """A device tensor with a Taichi-like surface.

    ``shape`` is the user-visible shape.  Vector fields over items
    (``items_dim``) are stored structure-of-arrays: user shape
    ``[..., n, 3]`` lives in storage ``[..., 3, n]`` so that per-step slices
    are coalesced for the kernels.  ``data`` is the storage tensor."""

    def __init__(self, shape, dtype, device, soa=False, needs_grad=True, name=""):
        self.shape = tuple(shape)
        self.soa = soa
        self.name = name
        store = self.shape
        if soa:
            store = self.shape[:-2] + (self.shape[-1], self.shape[-2])
        self.data = torch.zeros(store, dtype=dtype, device=device)
        self._grad = None
        self._needs_grad = needs_grad

    @property
    def view(self):
        return self.data.transpose(-1, -2) if self.soa else self.data

    @property
    def grad(self):
        if self._grad is None:
            if not self._needs_grad:
                raise AttributeError(f"field {self.name} has no gradient")
            g = Field.__new__(Field)
            g.shape, g.soa, g.name = self.shape, self.soa, self.name + ".grad"
            g.data = torch.zeros_like(self.data)
            g._grad, g._needs_grad = None, False
            self._grad = g
            _grad_fields.add(g)
        return self._grad

    def __getitem__(self, idx):
        if idx is None:
            return self.data.reshape(-1)[0].item()
        return self.view[idx]

    def __setitem__(self, idx, value):
        if idx is None:
            self.data.reshape(-1)[0] = value
            return
        self.view[idx] = torch.as_tensor(value, dtype=self.data.dtype, device=self.data.device)

    def fill(self, value):
        self.data.fill_(value)

    def to_numpy(self):
        return self.view.detach().cpu().numpy()

    def from_numpy(self, arr):
        self.view.copy_(torch.as_tensor(np.asarray(arr), dtype=self.data.dtype))

    def at(self, t):
        """Storage slice of leading index t (contiguous)."""
        return self.data[t]


def lazy_grad(*fields):
    """Allocate the gradient of each field (``ti.root.lazy_grad``)."""
    for f in fields:
        f.grad


class Tape:
    """``with Tape(loss): ...`` -- record forward kernels, replay adjoints."""

    def __init__(self, loss, clear_gradients=True):
        self.loss = loss
        self.clear_gradients = clear_gradients
        self.entries = []

    def __enter__(self):
        if self.clear_gradients:
            for g in list(_grad_fields):
                g.data.zero_()
        _tapes.append(self)
        return self

    def __exit__(self, exc_type, exc, tb):
        _tapes.pop()
        if exc_type is not None:
            return False
        self.loss.grad.fill(1.0)
        for adj in reversed(self.entries):
            adj()
        self.entries.clear()
        return False


def record(adjoint):
    """Register the adjoint closure of a forward kernel with the active tape."""
    if _tapes:
        _tapes[-1].entries.append(adjoint)


def recording():
    return bool(_tapes)
