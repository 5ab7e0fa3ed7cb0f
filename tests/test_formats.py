"""On-disk formats (photara_drender_amd.formats): bake / refine archives with the
reference's keys, COLMAP text models and pinhole world -> clip matrices.

The reference's own tests hold no COLMAP or archive fixture, so these checks
are pinned by what the mathematics fixes: the clip matrix against the pinhole
projection u = fx X/Z + cx written out per point, depth -1 / +1 at near / far,
the camera centre mapping to clip w = 0, quaternion rotations against scipy's
and against rotations about the axes, and sRGB decoding against its published
constants and continuity at the knee."""
import numpy as np
import pytest

from photara_drender_amd import formats as fm


def _camera(model="PINHOLE", w=640, h=480, params=(500.0, 520.0, 300.0, 250.0)):
    return fm.ColmapCamera(1, model, w, h, np.asarray(params, dtype=np.float64))


def _image(q=(0.9, 0.1, -0.3, 0.2), t=(0.1, -0.2, 4.0), name="a.png", image_id=1, camera_id=1):
    q = np.asarray(q, dtype=np.float64)
    return fm.ColmapImage(image_id, camera_id, name, fm.quaternion_to_rotation(q / np.linalg.norm(q)),
                          np.asarray(t, dtype=np.float64))


def test_quaternion_rotation_matches_scipy_and_axis_rotations():
    from scipy.spatial.transform import Rotation

    rng = np.random.default_rng(1)
    for _ in range(20):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        r = fm.quaternion_to_rotation(q)
        np.testing.assert_allclose(r, Rotation.from_quat([q[1], q[2], q[3], q[0]]).as_matrix(), atol=1e-12)
        np.testing.assert_allclose(r @ r.T, np.eye(3), atol=1e-12)
        assert abs(np.linalg.det(r) - 1.0) < 1e-12
    theta = 0.7
    c, s = np.cos(theta), np.sin(theta)
    rz = fm.quaternion_to_rotation((np.cos(theta / 2), 0.0, 0.0, np.sin(theta / 2)))
    np.testing.assert_allclose(rz, [[c, -s, 0], [s, c, 0], [0, 0, 1]], atol=1e-12)
    rx = fm.quaternion_to_rotation((np.cos(theta / 2), np.sin(theta / 2), 0.0, 0.0))
    np.testing.assert_allclose(rx, [[1, 0, 0], [0, c, -s], [0, s, c]], atol=1e-12)


@pytest.mark.parametrize("model,params", [("PINHOLE", (500.0, 520.0, 300.0, 250.0)),
                                          ("SIMPLE_PINHOLE", (610.0, 330.0, 210.0))])
def test_world_to_clip_is_the_pinhole_projection(model, params):
    cam, img = _camera(model, 640, 480, params), _image()
    fx, fy, cx, cy = (params[0], params[1], params[2], params[3]) if model == "PINHOLE" else \
        (params[0], params[0], params[1], params[2])
    near, far = 0.5, 20.0
    m = fm.world_to_clip(cam, img, near, far).astype(np.float64)
    rng = np.random.default_rng(2)
    centre = fm.camera_position(img)
    # points in front of the camera: camera-space coordinates mapped back to world space
    pc = np.column_stack([rng.uniform(-2, 2, 64), rng.uniform(-2, 2, 64), rng.uniform(1.0, 15.0, 64)])
    pw = (img.rotation.T @ (pc - img.translation).T).T
    clip = np.column_stack([pw, np.ones(64)]) @ m.T
    ndc = clip[:, :3] / clip[:, 3:4]
    u = fx * pc[:, 0] / pc[:, 2] + cx
    v = fy * pc[:, 1] / pc[:, 2] + cy
    np.testing.assert_allclose(ndc[:, 0], 2.0 * u / 640 - 1.0, rtol=0, atol=2e-5)
    np.testing.assert_allclose(ndc[:, 1], 2.0 * v / 480 - 1.0, rtol=0, atol=2e-5)
    np.testing.assert_allclose(clip[:, 3], pc[:, 2], rtol=1e-5)
    # depth: -1 at the near plane, +1 at the far plane, monotone between
    for depth, expect in ((near, -1.0), (far, 1.0)):
        p = img.rotation.T @ (np.array([0.3, -0.1, depth]) - img.translation)
        c = m @ np.r_[p, 1.0]
        assert abs(c[2] / c[3] - expect) < 1e-4
    assert np.all(np.diff(ndc[np.argsort(pc[:, 2]), 2]) > 0)
    # the camera centre sits on the w = 0 plane
    assert abs((m @ np.r_[centre, 1.0])[3]) < 1e-5


def test_pinhole_parameters_reject_distorted_models():
    with pytest.raises(ValueError, match="OPENCV"):
        fm.pinhole_parameters(_camera("OPENCV", params=(1, 1, 1, 1, 0, 0, 0, 0)))
    with pytest.raises(ValueError):
        fm.pinhole_parameters(_camera("PINHOLE", params=(1, 1, 1)))


def test_depth_range_rules():
    img = _image(q=(1, 0, 0, 0), t=(0, 0, 0))
    assert fm.depth_range(img) == (1e-3, 1e3)
    assert fm.depth_range(img, near=0.2, far=3.0) == (0.2, 3.0)
    mesh = np.array([[0, 0, 2.0], [0, 0, 8.0], [0, 0, -5.0]])
    assert fm.depth_range(img, mesh) == pytest.approx((1.0, 12.0))
    assert fm.depth_range(img, mesh, near=0.1) == pytest.approx((0.1, 12.0))
    with pytest.raises(ValueError, match="near and far"):  # near floors at 1e-4, above 1.5x the only depth
        fm.depth_range(img, np.array([[0, 0, 1e-5]]))
    with pytest.raises(ValueError, match="behind"):
        fm.depth_range(img, np.array([[0, 0, -1.0]]))
    with pytest.raises(ValueError):
        fm.depth_range(img, near=2.0, far=1.0)


def _write_model(root, images_block):
    (root / "cameras.txt").write_text("# Camera list\n# CAMERA_ID, MODEL, WIDTH, HEIGHT, PARAMS[]\n"
                                      "1 PINHOLE 8 6 10 11 4 3\n2 SIMPLE_PINHOLE 4 4 7 2 2\n")
    (root / "images.txt").write_text(images_block)


def test_read_colmap_text_model(tmp_path):
    _write_model(tmp_path, "# Image list\n"
                           "3 0 0 0 2 0.5 0 1 2 c.png\n1 2 3\n"
                           "1 2 0 0 0 0 0 3 1 a name with spaces.png\n10 20 -1 11 21 4\n")
    cams, imgs = fm.read_colmap_text_model(tmp_path)
    assert sorted(cams) == [1, 2]
    assert cams[1].model == "PINHOLE" and (cams[1].width, cams[1].height) == (8, 6)
    np.testing.assert_array_equal(cams[2].parameters, [7, 2, 2])
    assert [im.image_id for im in imgs] == [1, 3]
    assert imgs[0].name == "a name with spaces.png" and imgs[0].camera_id == 1
    np.testing.assert_allclose(imgs[0].rotation, np.eye(3))  # (2,0,0,0) normalised
    np.testing.assert_allclose(imgs[1].rotation, np.diag([-1.0, -1.0, 1.0]), atol=1e-12)  # 180 deg about z
    np.testing.assert_array_equal(imgs[1].translation, [0.5, 0, 1])


def test_read_colmap_text_model_errors(tmp_path):
    _write_model(tmp_path, "1 1 0 0 0 0 0 3 1 a.png\n")
    with pytest.raises(ValueError, match="points2D"):
        fm.read_colmap_text_model(tmp_path)
    _write_model(tmp_path, "1 0 0 0 0 0 0 3 1 a.png\n1 2 3\n")
    with pytest.raises(ValueError, match="quaternion"):
        fm.read_colmap_text_model(tmp_path)
    _write_model(tmp_path, "1 1 0 0 0 0 0 3\n1 2 3\n")
    with pytest.raises(ValueError, match="record"):
        fm.read_colmap_text_model(tmp_path)


def test_srgb_to_linear_constants():
    x = np.array([0.0, 0.04045, 0.5, 1.0], dtype=np.float32)
    np.testing.assert_allclose(fm.srgb_to_linear(x), [0.0, 0.04045 / 12.92, 0.21404114, 1.0], rtol=1e-5, atol=1e-7)
    # the two branches meet at the knee
    assert abs(((0.04045 + 0.055) / 1.055) ** 2.4 - 0.04045 / 12.92) < 1e-7


def test_load_colmap_projection(tmp_path):
    from PIL import Image

    _write_model(tmp_path, "2 1 0 0 0 0 0 3 1 b.png\nx\n1 1 0 0 0 0 0 3 1 a.png\nx\n"
                           "3 1 0 0 0 0 0 5 2 c.png\nx\n")
    pics = tmp_path / "pics"
    pics.mkdir()
    rng = np.random.default_rng(3)
    raw = {}
    for name, (w, h) in (("a.png", (8, 6)), ("b.png", (8, 6)), ("c.png", (4, 4))):
        raw[name] = rng.integers(0, 256, size=(h, w, 3), dtype=np.uint8)
        Image.fromarray(raw[name], "RGB").save(pics / name)
        Image.fromarray(raw[name][..., 0], "L").save(tmp_path / f"mask_{name}")
    mesh = np.array([[0, 0, -1.0], [0, 0, 1.0]])
    proj = fm.load_colmap_projection(tmp_path, pics, mesh_positions=mesh)
    assert proj.image_names == ("a.png", "b.png", "c.png")
    np.testing.assert_array_equal(proj.images[0], raw["a.png"].astype(np.float32) / np.float32(255.0))
    assert proj.world_to_clip.shape == (3, 4, 4) and proj.world_to_clip.dtype == np.float32
    np.testing.assert_allclose(proj.camera_positions, [[0, 0, -3], [0, 0, -3], [0, 0, -5]])
    # near / far from the mesh: depths 2..4 (a, b) -> 1, 6
    np.testing.assert_allclose(proj.world_to_clip[0], fm.world_to_clip(
        fm.read_colmap_text_model(tmp_path)[0][1], fm.read_colmap_text_model(tmp_path)[1][0], 1.0, 6.0))
    lin = fm.load_colmap_projection(tmp_path, pics, image_names=["b.png"], linearize_srgb=True)
    np.testing.assert_allclose(lin.images[0], fm.srgb_to_linear(raw["b.png"].astype(np.float32) / 255.0))
    assert fm.load_colmap_projection(tmp_path, pics, image_stride=2).image_names == ("a.png", "c.png")
    assert fm.load_colmap_projection(tmp_path, pics, max_images=1).image_names == ("a.png",)
    with pytest.raises(FileNotFoundError):
        fm.load_colmap_projection(tmp_path, pics, image_names=["zz.png"])
    with pytest.raises(ValueError):
        fm.load_colmap_projection(tmp_path, pics, image_stride=0)
    Image.fromarray(raw["b.png"][:4], "RGB").save(pics / "a.png")
    with pytest.raises(ValueError, match="expected 8x6"):
        fm.load_colmap_projection(tmp_path, pics)
    masks = fm.load_projection_masks(tmp_path, ["mask_c.png", "mask_b.png"])
    np.testing.assert_array_equal(masks[1], raw["b.png"][..., 0].astype(np.float32) / np.float32(255.0))


def test_bake_archive_round_trip(tmp_path):
    rng = np.random.default_rng(4)
    color = rng.random((16, 16, 4), dtype=np.float32)
    conf = rng.random((16, 16), dtype=np.float32)
    view = rng.integers(-1, 5, size=(16, 16)).astype(np.int32)
    valid = (conf > 0.5).astype(np.float32)
    path = fm.save_bake_npz(tmp_path / "bake.npz", color, conf, view, valid, texture_source="rgb",
                            source_images_path="/data/images", mask_applied=True, visibility_mode="ray_query")
    with np.load(path, allow_pickle=False) as raw:
        assert set(raw.files) == set(fm.BAKE_KEYS)
    back = fm.load_bake_npz(path)
    np.testing.assert_array_equal(back["color"], color)
    np.testing.assert_array_equal(back["source_view"], view)
    assert back["texture_source"] == "rgb" and back["mask_applied"] is True and back["used_ray_query"] is False
    assert back["visibility_mode"] == "ray_query" and back["color_space"] == "linear"
    pairs = rng.random((5, 2, 2), dtype=np.float32)
    hist = np.linspace(1, 0, 7, dtype=np.float32)
    path = fm.save_bake_npz(tmp_path / "sub" / "refined", color, conf, view, valid, history=hist, seam_pairs=pairs)
    assert path.name == "refined.npz"
    back = fm.load_bake_npz(path)
    assert set(back) == set(fm.BAKE_KEYS) | set(fm.REFINE_KEYS)
    assert back["optimized"] is True and back["seam_history"].shape == (0,)
    np.testing.assert_array_equal(back["seam_pairs"], pairs)
    np.testing.assert_array_equal(back["history"], hist)
    np.savez_compressed(tmp_path / "bad.npz", color=color)
    with pytest.raises(KeyError):
        fm.load_bake_npz(tmp_path / "bad.npz")
