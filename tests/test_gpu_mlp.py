"""The learned-shake MLP kernels (BF16 MFMA, hand-written forward and
backward) against a plain PyTorch FP32 reference of the same network.

Tolerances: the HIP path rounds weights, inputs and hidden activations to
BF16 (8-bit mantissa, relative step 2^-8 = 3.9e-3) and accumulates in FP32;
the forward reference applies the same roundings, so outputs agree to
~1e-2 relative.  The backward pass also carries BF16 gradients between
layers, so gradients are compared by relative Frobenius error <= 5e-2 per
tensor against FP32 autograd.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

K0, H = 512, 256


def _bf(t):
    return t.to(torch.bfloat16).float()


def ref_forward(x, p, lay, round_hidden=True):
    def w(name):
        o, s = lay[name]
        return torch.from_numpy(p[o:o + int(np.prod(s))].reshape(s)).cuda()
    h = x
    for i in (1, 2, 3):
        z = h @ _bf(w(f"W{i}")) + w(f"b{i}")
        h = torch.nn.functional.gelu(z)
        if round_hidden:
            h = _bf(h)
    return (h @ _bf(w("W4")) + w("b4"))[:, :4]


@pytest.fixture(scope="module")
def net(gpu_available):
    from pyopenlpt_amd.learned import ShakeNet
    n = ShakeNet(max_batch=1024, seed=3)
    yield n
    n.close()


def _batch(n, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    x = torch.zeros((n, K0), device="cuda")
    x[:, :484] = torch.randn((n, 484), device="cuda", generator=g)
    return x.to(torch.bfloat16).contiguous()


@pytest.mark.parametrize("n", [1, 63, 64, 200, 1024])
def test_forward_matches_torch(net, n):
    from pyopenlpt_amd.learned import param_layout
    x = _batch(n, 10 + n)
    y = net.infer(x)
    torch.cuda.synchronize()
    want = ref_forward(x.float(), net.params(), param_layout())
    err = (y - want).abs().max().item()
    scale = want.abs().max().item()
    assert err <= 2e-2 * max(scale, 1e-3), (err, scale)


def test_backward_matches_autograd(net):
    from pyopenlpt_amd.learned import param_layout
    lay = param_layout()
    p = net.params()
    n = 500
    x = _batch(n, 7)
    g = torch.Generator(device="cuda").manual_seed(8)
    t = torch.zeros((n, 4), device="cuda")
    t[:, :3] = 0.3 * torch.randn((n, 3), device="cuda", generator=g)
    t[:, 3] = (torch.rand(n, device="cuda", generator=g) < 0.4).float()
    loss = net.train_step(x, t.contiguous(), lr=0.0, ghost_weight=0.7)
    grads = net.grads()
    assert np.array_equal(net.params(), p)          # lr = 0 leaves parameters unchanged
    # FP32 autograd reference (weights rounded as the kernels use them)
    ws = {}
    for name, (o, s) in lay.items():
        a = torch.from_numpy(p[o:o + int(np.prod(s))].reshape(s)).cuda()
        ws[name] = (_bf(a) if name.startswith("W") else a).clone().requires_grad_(True)
    h = x.float()
    for i in (1, 2, 3):
        h = torch.nn.functional.gelu(h @ ws[f"W{i}"] + ws[f"b{i}"])
    y = (h @ ws["W4"] + ws["b4"])[:, :4]
    gl = t[:, 3]
    reg = ((y[:, :3] - t[:, :3]) ** 2).sum(1) * (1 - gl)
    bce = torch.nn.functional.binary_cross_entropy_with_logits(y[:, 3], gl, reduction="none")
    ref_loss = (reg + 0.7 * bce).mean()
    ref_loss.backward()
    assert abs(loss - ref_loss.item()) <= 2e-2 * abs(ref_loss.item())
    for name, (o, s) in lay.items():
        got = grads[o:o + int(np.prod(s))].reshape(s)
        want = ws[name].grad.cpu().numpy()
        if name == "W1":
            assert np.abs(got[484:]).max() == 0.0   # padded inputs get no gradient
            got, want = got[:484], want[:484]
        if name in ("W4", "b4"):
            assert np.abs(got[..., 4:]).max() == 0.0
            got, want = got[..., :4], want[..., :4]
        rel = np.linalg.norm(got - want) / max(np.linalg.norm(want), 1e-12)
        assert rel <= 5e-2, (name, rel)


def test_training_reduces_loss(gpu_available):
    """A learnable synthetic map (targets linear in the inputs, ghost label
    from the sign of another projection): 200 Adam steps cut the loss."""
    from pyopenlpt_amd.learned import ShakeNet
    net = ShakeNet(max_batch=512, seed=1)
    g = torch.Generator(device="cuda").manual_seed(5)
    A = torch.randn((484, 4), device="cuda", generator=g) / 22.0
    losses = []
    for step in range(200):
        x = _batch(512, 100 + step)
        proj = x.float()[:, :484] @ A
        t = torch.cat([0.5 * torch.tanh(proj[:, :3]), (proj[:, 3:] > 0).float()], 1).contiguous()
        losses.append(net.train_step(x, t, lr=1e-3, ghost_weight=0.5))
    net.close()
    assert np.mean(losses[-20:]) < 0.5 * np.mean(losses[:5]), (losses[:5], losses[-20:])
