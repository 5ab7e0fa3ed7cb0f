"""On-disk formats around the bake / refine path (host side, numpy only).

* The bake and refine archives of the reference's examples
  (``examples/bake_colmap.py:106-178``): ``np.savez_compressed`` files with
  the keys ``color, confidence, source_view, valid_mask, texture_source,
  source_images_path, mask_applied, visibility_mode, used_ray_query,
  color_space`` and, after refinement, ``optimized, history, seam_history,
  seam_pairs``.  Every value is a numeric or string array, so the files load
  with ``np.load(..., allow_pickle=False)``.
* COLMAP text models and the clip matrices the bake and the refiner take
  (``python/photara_drender/colmap.py:34-153``): ``cameras.txt`` /
  ``images.txt`` parsing, quaternion -> rotation, pinhole intrinsics, the
  near/far choice from the mesh and the OpenGL-style world -> clip matrix.
"""
from __future__ import annotations

from pathlib import Path
from typing import Dict, NamedTuple, Optional, Sequence, Tuple, Union

import numpy as np

BAKE_KEYS = ("color", "confidence", "source_view", "valid_mask", "texture_source", "source_images_path",
             "mask_applied", "visibility_mode", "used_ray_query", "color_space")
REFINE_KEYS = ("optimized", "history", "seam_history", "seam_pairs")


def save_bake_npz(path, color, confidence, source_view, valid_mask, *, texture_source="bake",
                  source_images_path="", mask_applied=False, visibility_mode="shadow_map", used_ray_query=False,
                  color_space="linear", history=None, seam_history=None, seam_pairs=None) -> Path:
    """Writes a bake (or, with ``history`` given, a refined bake) archive with the
    reference's keys (``examples/bake_colmap.py:109-120, 161-178``)."""
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    arrays = dict(
        color=np.asarray(color, dtype=np.float32),
        confidence=np.asarray(confidence, dtype=np.float32),
        source_view=np.asarray(source_view, dtype=np.int32),
        valid_mask=np.asarray(valid_mask, dtype=np.float32),
        texture_source=np.asarray(str(texture_source)),
        source_images_path=np.asarray(str(source_images_path)),
        mask_applied=np.asarray(bool(mask_applied)),
        visibility_mode=np.asarray(str(visibility_mode)),
        used_ray_query=np.asarray(bool(used_ray_query)),
        color_space=np.asarray(str(color_space)),
    )
    if history is not None:
        arrays.update(
            optimized=np.asarray(True),
            history=np.asarray(history, dtype=np.float32),
            seam_history=np.asarray(np.empty(0) if seam_history is None else seam_history, dtype=np.float32),
            seam_pairs=np.asarray(np.empty((0, 2, 2)) if seam_pairs is None else seam_pairs,
                                  dtype=np.float32).reshape(-1, 2, 2),
        )
    np.savez_compressed(path, **arrays)
    return path if path.suffix == ".npz" else path.with_name(path.name + ".npz")


def load_bake_npz(path) -> Dict[str, object]:
    """Reads a bake / refine archive (no pickled objects): arrays stay arrays,
    0-d string and bool entries become Python values."""
    out: Dict[str, object] = {}
    with np.load(path, allow_pickle=False) as data:
        missing = [k for k in BAKE_KEYS if k not in data.files]
        if missing:
            raise KeyError(f"bake archive lacks keys {missing}")
        for key in data.files:
            value = data[key]
            out[key] = value.item() if value.ndim == 0 else value
    return out


class ColmapCamera(NamedTuple):
    camera_id: int
    model: str
    width: int
    height: int
    parameters: np.ndarray


class ColmapImage(NamedTuple):
    image_id: int
    camera_id: int
    name: str
    rotation: np.ndarray     # world -> camera
    translation: np.ndarray  # camera = rotation @ world + translation


def quaternion_to_rotation(quaternion: Sequence[float]) -> np.ndarray:
    """Rotation matrix of a unit quaternion (w, x, y, z), COLMAP's convention
    (``colmap.py:39-48``)."""
    w, x, y, z = (float(v) for v in quaternion)
    return np.array([
        [1.0 - 2.0 * (y * y + z * z), 2.0 * (x * y - w * z), 2.0 * (x * z + w * y)],
        [2.0 * (x * y + w * z), 1.0 - 2.0 * (x * x + z * z), 2.0 * (y * z - w * x)],
        [2.0 * (x * z - w * y), 2.0 * (y * z + w * x), 1.0 - 2.0 * (x * x + y * y)],
    ])


def _records(path: Path):
    # non-empty, non-comment lines (colmap.py:34-36)
    with path.open("r", encoding="utf-8") as stream:
        return [s.strip() for s in stream if s.strip() and not s.startswith("#")]


def read_colmap_text_model(sparse_path: Union[str, Path]) -> Tuple[Dict[int, ColmapCamera], Tuple[ColmapImage, ...]]:
    """``cameras.txt`` and ``images.txt`` of a COLMAP text model (points ignored),
    images in id order (``colmap.py:51-90``, same records and errors)."""
    root = Path(sparse_path)
    cameras: Dict[int, ColmapCamera] = {}
    for rec in _records(root / "cameras.txt"):
        f = rec.split()
        cameras[int(f[0])] = ColmapCamera(int(f[0]), f[1], int(f[2]), int(f[3]), np.asarray(f[4:], dtype=np.float64))
    recs = _records(root / "images.txt")
    if len(recs) % 2:
        raise ValueError("COLMAP images.txt must contain one points2D line after each image line")
    images = []
    for rec in recs[0::2]:
        f = rec.split(maxsplit=9)
        if len(f) != 10:
            raise ValueError(f"invalid COLMAP image record: {rec[:120]}")
        q = np.asarray(f[1:5], dtype=np.float64)
        norm = float(np.linalg.norm(q))
        if not np.isfinite(norm) or norm <= 0.0:
            raise ValueError(f"image {f[0]} contains an invalid quaternion")
        images.append(ColmapImage(int(f[0]), int(f[8]), f[9], quaternion_to_rotation(q / norm),
                                  np.asarray(f[5:8], dtype=np.float64)))
    images.sort(key=lambda im: im.image_id)
    return cameras, tuple(images)


def pinhole_parameters(camera: ColmapCamera) -> Tuple[float, float, float, float]:
    """(fx, fy, cx, cy) of a PINHOLE or SIMPLE_PINHOLE camera (``colmap.py:93-101``)."""
    p = camera.parameters
    if camera.model == "PINHOLE" and p.size == 4:
        return float(p[0]), float(p[1]), float(p[2]), float(p[3])
    if camera.model == "SIMPLE_PINHOLE" and p.size == 3:
        return float(p[0]), float(p[0]), float(p[1]), float(p[2])
    raise ValueError(f"camera {camera.camera_id} uses {camera.model}; undistort it to PINHOLE or SIMPLE_PINHOLE "
                     "before projection")


def depth_range(image: ColmapImage, mesh_positions: Optional[np.ndarray] = None, near: Optional[float] = None,
                far: Optional[float] = None) -> Tuple[float, float]:
    """Near and far planes: as given, else from the mesh's camera depths (half the
    nearest, 1.5x the farthest), else 1e-3 / 1e3 (``colmap.py:104-128``)."""
    if near is not None and far is not None:
        if not 0.0 < near < far:
            raise ValueError("near and far must satisfy 0 < near < far")
        return float(near), float(far)
    if mesh_positions is None:
        n = 1e-3 if near is None else float(near)
        f = 1e3 if far is None else float(far)
    else:
        z = np.asarray(mesh_positions, dtype=np.float64) @ image.rotation[2] + image.translation[2]
        z = z[z > 1e-6]
        if z.size == 0:
            raise ValueError(f"mesh is entirely behind COLMAP camera for {image.name}")
        n = max(1e-4, float(z.min()) * 0.5) if near is None else float(near)
        f = float(z.max()) * 1.5 if far is None else float(far)
    if not 0.0 < n < f:
        raise ValueError("near and far must satisfy 0 < near < far")
    return n, f


def world_to_clip(camera: ColmapCamera, image: ColmapImage, near: float, far: float) -> np.ndarray:
    """float32 [4, 4] world -> clip matrix of a pinhole view (``colmap.py:131-153``):
    NDC x = 2 (fx X/Z + cx) / W - 1, y likewise, z = -1 at Z = near and +1 at Z = far."""
    fx, fy, cx, cy = pinhole_parameters(camera)
    R, t = image.rotation, image.translation
    sx, sy = 2.0 * fx / camera.width, 2.0 * fy / camera.height
    ox, oy = 2.0 * cx / camera.width - 1.0, 2.0 * cy / camera.height - 1.0
    a, b = (far + near) / (far - near), -2.0 * far * near / (far - near)
    m = np.zeros((4, 4))
    m[0, :3], m[0, 3] = sx * R[0] + ox * R[2], sx * t[0] + ox * t[2]
    m[1, :3], m[1, 3] = sy * R[1] + oy * R[2], sy * t[1] + oy * t[2]
    m[2, :3], m[2, 3] = a * R[2], a * t[2] + b
    m[3, :3], m[3, 3] = R[2], t[2]
    return m.astype(np.float32)


def camera_position(image: ColmapImage) -> np.ndarray:
    """World position of the camera centre: -R^T t."""
    return -(image.rotation.T @ image.translation)


def srgb_to_linear(image: np.ndarray) -> np.ndarray:
    """IEC 61966-2-1 decoding (``colmap.py:156-157``)."""
    image = np.asarray(image, dtype=np.float32)
    return np.where(image <= 0.04045, image / 12.92, ((image + 0.055) / 1.055) ** 2.4).astype(np.float32)


class ColmapProjection(NamedTuple):
    images: Tuple[np.ndarray, ...]   # [H, W, 3] float32 each
    world_to_clip: np.ndarray        # [V, 4, 4] float32, row-major
    camera_positions: np.ndarray     # [V, 3] float32
    image_names: Tuple[str, ...]


def _load_picture(path: Path, mode: str) -> np.ndarray:
    from PIL import Image

    with Image.open(path) as source:
        return np.asarray(source.convert(mode), dtype=np.float32) / np.float32(255.0)


def load_colmap_projection(sparse_path, images_path, *, mesh_positions=None, image_names=None, image_stride=1,
                           max_images=None, near=None, far=None, linearize_srgb=False) -> ColmapProjection:
    """Photographs and clip matrices of a COLMAP text model, with the selection
    rules and errors of ``colmap.py:192-266``: by name, then every
    ``image_stride``-th, then the first ``max_images``."""
    if image_stride < 1:
        raise ValueError("image_stride must be positive")
    cameras, images = read_colmap_text_model(sparse_path)
    wanted = None if image_names is None else set(image_names)
    chosen = [im for im in images if wanted is None or im.name in wanted]
    if wanted is not None:
        missing = wanted.difference(im.name for im in chosen)
        if missing:
            raise FileNotFoundError(f"images are absent from COLMAP model: {sorted(missing)}")
    chosen = chosen[::image_stride]
    if max_images is not None:
        chosen = chosen[:max_images]
    if not chosen:
        raise ValueError("no COLMAP images matched the selection")
    mesh = None
    if mesh_positions is not None:
        mesh = np.asarray(mesh_positions, dtype=np.float64)
        if mesh.ndim != 2 or mesh.shape[1] != 3:
            raise ValueError("mesh_positions must have shape [vertex_count, 3]")
        if mesh.shape[0] == 0 or not np.isfinite(mesh).all():
            raise ValueError("mesh_positions must be non-empty and finite")
    pictures, clips, centres = [], [], []
    for image in chosen:
        camera = cameras.get(image.camera_id)
        if camera is None:
            raise ValueError(f"image {image.name} references missing camera {image.camera_id}")
        picture = _load_picture(Path(images_path) / image.name, "RGB")
        if picture.shape[:2] != (camera.height, camera.width):
            raise ValueError(f"{image.name} has size {picture.shape[1]}x{picture.shape[0]}, expected "
                             f"{camera.width}x{camera.height}")
        if linearize_srgb:
            picture = srgb_to_linear(picture)
        n, f = depth_range(image, mesh, near, far)
        pictures.append(np.ascontiguousarray(picture))
        clips.append(world_to_clip(camera, image, n, f))
        centres.append(camera_position(image).astype(np.float32))
    return ColmapProjection(tuple(pictures), np.ascontiguousarray(clips, dtype=np.float32),
                            np.ascontiguousarray(centres, dtype=np.float32), tuple(im.name for im in chosen))


def load_projection_masks(masks_path, image_names: Sequence[str]) -> Tuple[np.ndarray, ...]:
    """Grayscale masks in [0, 1], one per name, in the given order (``colmap.py:269-285``)."""
    return tuple(np.ascontiguousarray(_load_picture(Path(masks_path) / name, "L")) for name in image_names)
