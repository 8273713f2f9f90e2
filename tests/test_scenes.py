"""Seeded scene generators: sizes, determinism and camera conventions."""
import numpy as np

from softrast_amd import scenes


def test_cube_grid_config():
    a = scenes.cube_grid_scene()
    b = scenes.cube_grid_scene()
    assert (a.width, a.height) == (1280, 720)
    assert a.num_tris == 12288 and len(a.draws[0].vertices) == 24576
    assert a.draws[0].indices.dtype == np.uint16
    assert np.array_equal(a.draws[0].vertices, b.draws[0].vertices)
    assert np.array_equal(a.textures[0], b.textures[0])
    assert a.textures[0].shape == (512, 512, 4) and np.all(a.textures[0][..., 3] == 255)


def test_spheres_config_small():
    s = scenes.spheres_scene(nspheres=40, ndraws=4, tex_size=64)
    assert s.num_tris == 40 * 250
    assert len(s.draws) == 4 and all(d.indices.dtype == np.uint32 for d in s.draws)
    for d in s.draws:
        assert d.indices.max() < len(d.vertices)
    assert len(s.draws[0].vertices) == 10 * 182


def test_reverse_z_projection():
    m = scenes.camera_mvp((0, 0, 0), (0, 0, 1), 60.0, 16 / 9).reshape(4, 4).T.astype(np.float64)
    def ndc_z(z):
        c = m @ np.array([0, 0, z, 1.0])
        return c[2] / c[3]
    assert abs(ndc_z(0.1) - 1.0) < 1e-5       # far parameter 0.1 -> depth 1
    assert abs(ndc_z(10000.0)) < 1e-5        # near parameter 10000 -> depth 0
    assert ndc_z(1.0) > ndc_z(2.0) > 0        # nearer is larger


def test_value_noise_deterministic_and_seeded():
    a = scenes.value_noise_rgba8(1, 64, 64)
    b = scenes.value_noise_rgba8(1, 64, 64)
    c = scenes.value_noise_rgba8(2, 64, 64)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a[..., :3].std() > 5
