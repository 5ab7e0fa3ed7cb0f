"""The textured raster chain (rasterize -> interpolate -> texture, forward and
backward through the C ABI) captured in a HIP graph: every replay, with eager
chains and synchronisations between the replays, gives the eager chain's
results - bit-identical, except the texel and uv-attribute gradients, which
are float-atomic sums and agree to 1e-4 relative.  The index buffer rewritten
in place between replays is picked up by the replay (the vertex adjacency is
rebuilt on the device).
"""
import numpy as np
import pytest

from photara_drender_amd import scenes

pytestmark = pytest.mark.gpu

ATOMIC = ("g_tex", "g_uv")


def same(got, ref, what):
    for k, r in ref.items():
        if k in ATOMIC:
            np.testing.assert_allclose(got[k], r, rtol=1e-4, atol=1e-6 * max(np.abs(r).max(), 1.0),
                                       err_msg=f"{what}: {k}")
        else:
            np.testing.assert_array_equal(got[k], r, err_msg=f"{what}: {k}")


def test_graph_replays_match_the_eager_chain():
    import torch

    from photara_drender_amd import _native, workloads

    sc = scenes.chain_scene(level=4, res=(256, 192), tex_size=256, seed=9)
    ctx = _native.Context(0, flags=0)
    chain = workloads.TexturedChain(ctx, sc)
    s = torch.cuda.Stream()
    chain.step(s)
    torch.cuda.synchronize()
    eager = chain.outputs()
    assert (eager["raster"][..., 3] != 0).mean() > 0.2
    graph = chain.capture(s)
    for r in range(3):
        for k in chain.NAMES:  # stale outputs must be overwritten by the replay
            getattr(chain, k).fill_(float("nan"))
        graph.replay()
        torch.cuda.synchronize()
        same(chain.outputs(), eager, f"replay {r}")
        chain.step(s)
        torch.cuda.synchronize()
    # the same index buffer rewritten in place: the replay sees the new mesh
    rotated = np.roll(sc.indices, 1, axis=1)
    with torch.cuda.stream(s):
        chain.idx.copy_(torch.tensor(rotated.view(np.int32)))
    graph.replay()
    torch.cuda.synchronize()
    replayed = chain.outputs()
    chain.step(s)
    torch.cuda.synchronize()
    fresh = chain.outputs()
    same(replayed, fresh, "replay after the indices changed")
    assert not np.array_equal(fresh["raster"], eager["raster"])  # rotated corners: other barycentrics
