"""The seeded scene generators: the denoiser's random training scenes can
leave the large-mesh benchmark scene's family (displaced icospheres) out."""
import numpy as np
import pytest

import hpt_scenes


def test_random_scene_layout_filter():
    sizes_all, sizes_kept = [], []
    for seed in range(2000, 2040):
        full, _ = hpt_scenes.random_scene(seed, 32, 32)
        kept, _ = hpt_scenes.random_scene(seed, 32, 32, layouts=(0, 1))
        sizes_all.append(max(p.vertices.shape[0] for p in full.primitives))
        sizes_kept.append(max(p.vertices.shape[0] for p in kept.primitives))
    # icospheres of level >= 4 have >= 2562 vertices; UV spheres 561, boxes and quads fewer
    assert max(sizes_kept) < 2562 <= max(sizes_all)
    # the default draw is the three-layout one, seed for seed
    a, _ = hpt_scenes.random_scene(2001, 32, 32)
    b, _ = hpt_scenes.random_scene(2001, 32, 32, layouts=(2, 1, 0))
    assert a.triangle_count == b.triangle_count
    with pytest.raises(ValueError):
        hpt_scenes.random_scene(1, 32, 32, layouts=())
