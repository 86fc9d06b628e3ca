"""Host tables of the batched bundle adjustment: the library's native builder (ct_ba_batch_tables) gives,
problem by problem, the tables of the per-problem builder (ba.build_tables) that the per-problem plans use:
the same observation order and CSR tables, and each camera-pair block holds the same observation pairs
(the native builder lists them in ascending (a, b))."""
from __future__ import annotations

import numpy as np

from ceres_mi355x.ba import build_tables
from ceres_mi355x.ba_batch import build_batch_tables
from oracle import generators as gen


def _problems():
    rng = np.random.default_rng(3)
    out = []
    for C, P in ((2, 50), (7, 300), (21, 400), (1, 20), (13, 90)):
        out.append(gen.make_bundle_adjustment(num_cameras=C, num_points=P, obs_per_point=(1, min(6, C)),
                                              seed=int(rng.integers(1 << 30))))
    # a point seen twice by the same camera and a point seen by no camera
    d = out[1]
    oc = np.append(d.obs_camera, [d.obs_camera[0]])
    op = np.append(d.obs_point, [d.obs_point[0]])
    out.append((oc, op, d.num_cameras, d.num_points + 1))
    return [(p.obs_camera, p.obs_point, p.num_cameras, p.num_points) if not isinstance(p, tuple) else p
            for p in out]


def test_batch_tables_equal_the_per_problem_tables():
    probs = _problems()
    tb = build_batch_tables([p[2] for p in probs], [p[3] for p in probs], [p[0] for p in probs],
                            [p[1] for p in probs])
    for k, (oc, op, C, P) in enumerate(probs):
        t = build_tables(oc, op, C, P)
        N, T, B = len(oc), t.block_terms.shape[0], C * (C + 1) // 2
        o0, t0 = tb["obase"][k], tb["tbase"][k]
        np.testing.assert_array_equal(tb["order"][o0:o0 + N], t.order)
        np.testing.assert_array_equal(tb["obs_camera"][o0:o0 + N], t.obs_camera)
        np.testing.assert_array_equal(tb["obs_point"][o0:o0 + N], t.obs_point)
        np.testing.assert_array_equal(tb["camera_obs"][o0:o0 + N], t.camera_obs)
        pb = tb["poff_base"][k]
        np.testing.assert_array_equal(tb["point_offsets"][pb:pb + P + 1], t.point_offsets)
        cb = tb["coff_base"][k]
        np.testing.assert_array_equal(tb["camera_offsets"][cb:cb + C + 1], t.camera_offsets)
        bb = tb["boff_base"][k]
        np.testing.assert_array_equal(tb["block_offsets"][bb:bb + B + 1], t.block_offsets)
        assert tb["num_terms"][k] == T
        mine = tb["block_terms"][2 * t0:2 * (t0 + T)].reshape(-1, 2)
        for blk in range(B):
            s, e = t.block_offsets[blk], t.block_offsets[blk + 1]
            ref = t.block_terms[s:e]
            ref = ref[np.lexsort((ref[:, 1], ref[:, 0]))]
            np.testing.assert_array_equal(mine[s:e], ref)


def test_batch_tables_refuse_bad_indices():
    import pytest

    from ceres_mi355x._native import NativeLibraryError

    with pytest.raises(NativeLibraryError):
        build_batch_tables([2], [3], [np.array([0, 2])], [np.array([0, 1])])
    with pytest.raises(NativeLibraryError):
        build_batch_tables([22], [3], [np.array([0])], [np.array([0])])
