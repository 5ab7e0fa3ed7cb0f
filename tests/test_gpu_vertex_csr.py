"""The rasterize backward's vertex -> corner adjacency, kept in the context
across calls: it is rebuilt on the device whenever the index buffer's words, T
or V differ from the ones it was built from, through the one-launch path
(V < 16,384) and the multi-kernel path (larger meshes).

One context runs a sequence of backward calls that each change one of those
things (the same buffer rewritten in place, a single word changed, a new
buffer with the same words, a shorter T over the same prefix, a larger V), and
every call's vertex gradients are bit-identical to the oracle's
(rasterize_backward.hlsl + reduce_vertex_gradients.hlsl restated in C).
"""
import numpy as np
import pytest

import oracle
from photara_drender_amd import scenes

pytestmark = pytest.mark.gpu


def bits_equal(a, b, what):
    a, b = np.ascontiguousarray(a, np.float32), np.ascontiguousarray(b, np.float32)
    assert a.shape == b.shape, what
    diff = a.view(np.uint32) != b.view(np.uint32)
    assert not diff.any(), f"{what}: {int(diff.sum())} values differ"


class Backward:
    """pdr_rasterize_backward over device tensors on one context."""

    def __init__(self):
        import torch

        from photara_drender_amd import _native

        self.torch, self.native = torch, _native
        self.ctx = _native.Context(0, flags=0)
        self.dev = torch.device("cuda:0")

    def __call__(self, pos, idx_dev, T, res, what):
        torch, native = self.torch, self.native
        H, W = res
        idx_host = idx_dev[:T].cpu().numpy().view(np.uint32)
        raster, _, _ = oracle.rasterize_forward(pos, idx_host, res, 0, False)
        grad = np.random.default_rng(T).standard_normal(raster.shape).astype(np.float32)
        ref = oracle.rasterize_backward(pos, idx_host, raster, grad)
        d_pos = torch.tensor(pos, device=self.dev)
        d_r = torch.tensor(raster, device=self.dev)
        d_g = torch.tensor(grad, device=self.dev)
        out = torch.full((pos.shape[0], 4), float("nan"), device=self.dev)
        st = torch.cuda.current_stream().cuda_stream
        native.check(native.lib().pdr_rasterize_backward(self.ctx.handle, d_pos.data_ptr(), pos.shape[0],
                                                         idx_dev.data_ptr(), T, d_r.data_ptr(), W, H, d_g.data_ptr(),
                                                         None, out.data_ptr(), st))
        bits_equal(out.cpu().numpy(), ref, what)


@pytest.mark.parametrize("level,res", [(4, (256, 256)), (6, (512, 512))], ids=["one_launch", "multi_kernel"])
def test_adjacency_follows_index_changes(level, res):
    import torch

    sc = scenes.chain_scene(level=level, res=res, tex_size=4, seed=5)
    pos, idx = sc.clip, sc.indices
    V, T = pos.shape[0], idx.shape[0]
    assert (V < 16384) == (level == 4)
    run = Backward()
    dev = run.dev
    buf = torch.tensor(idx.view(np.int32), device=dev)
    run(pos, buf, T, res, "first call")
    run(pos, buf, T, res, "same buffer again")
    # the same buffer rewritten in place: every triangle's corners rotated
    buf.copy_(torch.tensor(np.roll(idx, 1, axis=1).view(np.int32), device=dev))
    run(pos, buf, T, res, "rotated corners in place")
    # one word changed (a triangle takes a neighbour's vertex)
    one = np.roll(idx, 1, axis=1).copy()
    one[T // 2, 0] = one[T // 2 + 1, 0]
    buf.copy_(torch.tensor(one.view(np.int32), device=dev))
    run(pos, buf, T, res, "one word changed")
    # a new buffer holding the original words
    other = torch.tensor(idx.view(np.int32), device=dev)
    run(pos, other, T, res, "new buffer, original indices")
    # fewer triangles over the same prefix
    run(pos, other, T - 37, res, "shorter T")
    # more vertices (unused ones appended), same indices
    extra = np.concatenate([pos, np.tile(pos[:1], (5, 1))]).astype(np.float32)
    run(extra, other, T, res, "larger V")
    run(pos, other, T, res, "back to the original")
