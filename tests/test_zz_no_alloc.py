"""dl_mlp_backward_static (library-owned workspace, allocates nothing) against
dl_mlp_backward (caller workspace) through the C ABI: two calls in a row on
the same GPU, the second with a smaller batch, and the size limit."""

import sys

import pytest
import torch

from diffloco import _native as nat

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _inputs(B, I, H, O, dt, seed):
    g = torch.Generator().manual_seed(seed)
    r = lambda *s: torch.randn(*s, dtype=torch.float64, generator=g)
    W = [r(H, I) / I ** 0.5, 0.1 * r(H), r(O, H) / H ** 0.5, 0.1 * r(O)]
    return ([r(B, I).to("cuda", dt)] + [w.to("cuda", dt) for w in W] + [r(B, O).to("cuda", dt)])


def _grads(B, I, H, O, dt):
    z = lambda *s: torch.zeros(*s, dtype=dt, device="cuda")
    return [z(B, I), z(H, I), z(H), z(O, H), z(O)]


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_static_workspace_matches_caller_workspace(prec):
    dt = torch.float32 if prec == "f32" else torch.float64
    code = nat.dtype_code(dt)
    I, H, O = 179, 64, 4
    for call, B in enumerate((1024, 37)):     # the second call smaller than the first
        _say(f"mlp_backward_static {prec} call {call} B={B}")
        inp, W1, b1, W2, b2, og = _inputs(B, I, H, O, dt, seed=100 + call)
        ref = _grads(B, I, H, O, dt)
        got = _grads(B, I, H, O, dt)
        work = torch.zeros(B * (2 * H + O), dtype=dt, device="cuda")
        P = nat.ptr
        for act in (nat.ACT_TANH, nat.ACT_SIN):
            nat.call("dl_mlp_backward", code, act, B, I, H, O, P(inp), P(W1), P(b1), P(W2), P(b2),
                     P(og), *map(P, ref), P(work), nat.stream())
            nat.call("dl_mlp_backward_static", code, act, B, I, H, O, P(inp), P(W1), P(b1),
                     P(W2), P(b2), P(og), *map(P, got), nat.stream())
        torch.cuda.synchronize()
        for a, b in zip(got, ref):
            assert torch.isfinite(a).all()
            assert torch.equal(a, b)      # same kernels, same order: bit for bit
        assert ref[1].abs().max() > 0


def test_static_workspace_limit():
    _say("mlp_backward_static limit")
    I, H, O = 16, 64, 4
    B = (4 << 20) // (4 * (2 * H + O)) + 1        # one robot more than the FP32 limit
    z = torch.zeros(1, device="cuda")
    P = nat.ptr(z)
    rc = nat.lib().dl_mlp_backward_static(nat.F32, nat.ACT_TANH, B, I, H, O, P, P, P, P, P, P, P,
                                          P, P, P, P, nat.stream())
    assert rc == -1
    assert "static limit" in nat.last_error()
