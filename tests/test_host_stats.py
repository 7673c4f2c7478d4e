"""Host-side scene statistics on a host-only context (no GPU): Bvh.GetStats (bvh.cs:79-125) of mesh
and scene BVHs against the oracle's restatement, and the build counters of the Counters.Type book
(mesh.cs:38-45, scene.cs:528-540)."""
import numpy as np
import pytest


def _scene_spec():
    from sharpray_amd import scenes
    return scenes.config_plain_cpu(scale=0.125, spp=1, level=3)


def test_bvh_stats_match_oracle(oracle_lib):
    from sharpray_amd import scenes
    spec = _scene_spec()
    scene, _ = scenes.instantiate(spec, device=-1)
    scene.build()
    osc = oracle_lib.OracleScene(spec)
    for m in [None] + list(range(len(spec.meshes))):
        g = scene.bvh_stats(m)
        o = osc.bvh_stats(m)
        assert g["NodeCount"] == o["nodeCount"] and g["DepthMax"] == o["depthMax"]
        assert g["LeafCount"] == o["leafCount"]
        assert (g["LeafSizeMin"], g["LeafSizeMax"]) == (o["leafSizeMin"], o["leafSizeMax"])
        assert np.float32(g["SahCost"]) == np.float32(o["sah"]), (m, g["SahCost"], o["sah"])
        # bvh.cs:90-91: the averages from the record's sums
        assert g["LeafSizeAvg"] == pytest.approx(o["shapeCount"] / o["leafCount"], rel=1e-6)
        assert g["LeafDepthAvg"] == pytest.approx(o["depthWeighted"] / o["shapeCount"], rel=1e-6)
        nodes, _ = scene.bvh(m)
        assert g["NodeCount"] == nodes.shape[0]


def test_bvh_stats_closed_forms():
    """(a) Two triangles: the SAH never splits them (leaf 0.5 x 2 x SA = SA <= the split's traverse
    cost alone), so one root leaf of 2 at depth 0 with SahCost 1. (b) Two far clusters of 4
    identical unit triangles: a root and two leaves of 4 (identical centroids cannot split); SahCost
    = 1 + 2 x 0.5 x 4 x (leaf area 2 / root area 22), summed in the reference's order."""
    from sharpray_amd import Mesh, Scene, triangles_build
    sc = Scene(-1)
    t = [0, 0, 0, 1, 0, 0, 0, 1, 0]
    far = [10, 0, 0, 11, 0, 0, 10, 1, 0]
    mid = sc.mesh_id(Mesh(triangles_build(np.array([t, far], np.float32), None, None), split_bins=8))
    st = sc.bvh_stats(mid)
    assert (st["NodeCount"], st["DepthMax"], st["LeafCount"], st["LeafSizeMin"], st["LeafSizeMax"]) == (1, 0, 1, 2, 2)
    assert (st["LeafSizeAvg"], st["LeafDepthAvg"], st["SahCost"]) == (2.0, 0.0, 1.0)
    mid = sc.mesh_id(Mesh(triangles_build(np.array([t] * 4 + [far] * 4, np.float32), None, None), split_bins=8))
    st = sc.bvh_stats(mid)
    assert (st["NodeCount"], st["DepthMax"], st["LeafCount"], st["LeafSizeMin"], st["LeafSizeMax"]) == (3, 1, 2, 4, 4)
    assert (st["LeafSizeAvg"], st["LeafDepthAvg"]) == (4.0, 1.0)
    f = np.float32
    ratio = f(2) / f(22)
    leaf = f(f(0.5) * ratio) * f(4)
    assert f(st["SahCost"]) == f(f(leaf + leaf) + f(1) * f(1))


def test_build_counters_in_counter_book():
    """With counters on, mesh creation and scene build bump MeshTriangle, SceneObject,
    SceneBvhNodes and SceneBvhDepth as the reference does when handed a Counters object; the
    times are non-negative; the render-path entries stay unset on a host-only context."""
    from sharpray_amd import _native as N, scenes
    spec = _scene_spec()
    sc2, _ = scenes.instantiate(spec, device=-1)  # reference numbers without counters
    sc2.build()
    assert sc2.counters_all() == {}
    scene, _ = scenes.instantiate(spec, device=-1, counters=True)
    scene.build()
    c = scene.counters_all()
    assert c["MeshTriangle"] == spec.triangle_count
    assert c["SceneObject"] == len(spec.objects)
    st = sc2.bvh_stats()
    assert c["SceneBvhNodes"] == st["NodeCount"] and c["SceneBvhDepth"] == st["DepthMax"]
    assert c["TimeMeshBvhBuild"] >= 0 and c["TimeSceneBvhBuild"] >= 0
    assert "Pixel" not in c and "SampleCountMin" not in c
    full = scene.counters_all(include_unset=True)
    assert len(full) == N.SR_COUNTER_TYPE_COUNT == 51
    assert full["SampleCountMin"] == 2 ** 63 - 1 and full["SampleCountMax"] == -2 ** 63
    scene.build()  # idempotent: no second bump (scene.cs:519-524)
    assert scene.counters_all()["SceneObject"] == len(spec.objects)
    scene.counters_reset()
    assert scene.counters_all() == {}


def test_cpu_configuration_leaf_sizes_under_reference_sah():
    """The CPU configuration asks for at most 4 triangles per leaf, a rule the reference's builder
    does not have (its SAH leaf test decides, bvh.cs:332,472-476). The build keeps the reference's
    rule, so node order stays bit-exact, and this is the leaf-size report that takes the rule's
    place: 16 bins over the level-5 icosphere give leaves of 1-6 triangles, 2.88 on average, and
    all but 14 of 7120 leaves hold at most 4."""
    from sharpray_amd import scenes
    spec = scenes.config_plain_cpu()
    scene, _ = scenes.instantiate(spec, device=-1)
    st = scene.bvh_stats(0)
    nodes, _ = scene.bvh(0)
    counts = nodes[:, 7].view(np.int32)
    leaves = counts[counts > 0]
    assert st["LeafCount"] == leaves.size == 7120
    assert (st["LeafSizeMin"], st["LeafSizeMax"]) == (1, 6)
    assert abs(st["LeafSizeAvg"] - 2.8764) < 1e-3
    assert int((leaves > 4).sum()) == 14
