"""The tile autotuner: times whole steps for tiles 8, 16 and 32, keeps the
fastest, remembers it in a JSON cache, and the chosen tile gives the same
raster as any other (tiles change the work split, never the result)."""
import json

import numpy as np
import pytest
import torch

from diffsoup_amd import scenes

pytestmark = pytest.mark.gpu


def test_autotune_picks_a_tile_and_caches_it(dev, tmp_path):
    from diffsoup_amd import fused
    s = scenes.sphere_torus_soup(8, n_surface=3000, n_float=800, views=3, res=(200, 260))
    prm = fused.SoupParams.from_soup(s, dev)
    mvp = torch.from_numpy(s.mvp).to(dev)
    target = torch.from_numpy(s.target).to(dev)
    mk = lambda tile: fused.FusedStep(s.resolutions, s.T, s.verts.shape[0], s.alpha_level, s.levels, 8, 32,
                                      tile=tile, device=dev)
    cache = str(tmp_path / "tiles.json")
    st = mk(16)
    times = st.autotune(prm, mvp, target, reps=2, cache=cache)
    assert set(times) == {8, 16, 32} and all(t is not None and t > 0 for t in times.values())
    assert st.tile == min(times, key=times.get)
    assert json.load(open(cache))[st.tune_key()] == st.tile
    st.run(prm, mvp, target, seed=31)
    ref = mk(16)
    ref.run(prm, mvp, target, seed=31)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(st.rast.cpu().numpy().view(np.uint32), ref.rast.cpu().numpy().view(np.uint32))
    # a second step object with the same configuration reads the cache
    st2 = mk(8)
    assert st2.autotune(prm, mvp, target, reps=2, cache=cache) == {st.tile: None}
    assert st2.tile == st.tile
