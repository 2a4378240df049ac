"""The emulator's torch operators without a GPU: both register with their
schemas, and their fake-tensor rules and the registered autograd formula
give the shapes and dtypes of the network's output and of every parameter
gradient (FakeTensorMode runs no kernel)."""
import torch
from torch._subclasses.fake_tensor import FakeTensorMode


def test_ops_register_and_trace():
    from exoplasim_amd import radiation as r

    f = r.emulator_op()
    fwd = torch.ops.exoplasim_amd.lw_mlp_forward.default._schema
    bwd = torch.ops.exoplasim_amd.lw_mlp_backward.default._schema
    assert len(fwd.arguments) == 9 and len(fwd.returns) == 8
    assert len(bwd.arguments) == 12 and len(bwd.returns) == 8
    with FakeTensorMode():
        x = torch.empty(96, r.NIN)
        ps = []
        for nin, nout in r.LAYERS:
            ps += [torch.empty(nout, nin, requires_grad=True),
                   torch.empty(nout, requires_grad=True)]
        y = f(x, *ps)
        assert y.shape == (96, r.NOUT) and y.dtype == torch.float32
        y.sum().backward()
        for p in ps:
            assert p.grad is not None and p.grad.shape == p.shape
