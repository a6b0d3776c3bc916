"""reg_drot.autograd.unpack: the per-problem views of a packed buffer and the
gradient they send back (on the CPU; the packed layout is plain slicing)."""

import pytest

torch = pytest.importorskip("torch")


class _Layout:
    """The slice of PackedBatch that unpack uses: problem i is an
    [n_cols, ld] block at offset off_i, of which the first n_rows rows count."""

    def __init__(self, shapes, pad=4):
        self.shapes = shapes
        self.off, self.ld = [], []
        o = 0
        for n, m in shapes:
            ld = -(-n // pad) * pad
            self.off.append(o)
            self.ld.append(ld)
            o += ld * m
        self.total = o

    def block(self, t, i):
        n, m = self.shapes[i]
        return t[self.off[i]: self.off[i] + self.ld[i] * m].view(m, self.ld[i])[:, :n]


def test_unpack_views_and_gradient():
    from reg_drot import autograd
    g = torch.Generator().manual_seed(0)
    pb = _Layout([(5, 3), (8, 2), (3, 7), (1, 1)])
    xp = torch.randn(pb.total, generator=g, requires_grad=True)
    ws = [torch.randn(m, n, generator=g) for n, m in pb.shapes]
    views = autograd.unpack(xp, pb)
    for i, v in enumerate(views):
        assert torch.equal(v, pb.block(xp.detach(), i))
    loss = sum((w * v).sum() for w, v in zip(ws, views))
    loss.backward()
    want = torch.zeros(pb.total)
    for i, w in enumerate(ws):
        pb.block(want, i).copy_(w)
    assert torch.equal(xp.grad, want)  # padding rows get zero


def test_unpack_unused_views_get_no_gradient():
    from reg_drot import autograd
    pb = _Layout([(4, 4), (4, 4)])
    xp = torch.ones(pb.total, requires_grad=True)
    views = autograd.unpack(xp, pb)
    views[1].sum().backward()
    assert float(xp.grad[: 16].abs().sum()) == 0.0
    assert float(xp.grad[16:].sum()) == 16.0
