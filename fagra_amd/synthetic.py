"""Seeded synthetic bundle-adjustment problems in the BAL camera model.

Cameras are ``[angle-axis(3), t(3), f, k1, k2]`` and project with
``P = R X + t``, ``q = -P.xy / P.z``, ``pixel = f (1 + k1 r2 + k2 r2^2) q``
(examples/bal/model.rs:147-156).  Cameras sit on horizontal rings around the
origin and look at it along their -z axis (the BAL convention,
docs/bal-performance.md:31-37); points are uniform in [-2, 2]^3.

Workloads (names carry their sizes):

* ``ring16``        16 cameras, 2,000 points, tracks U{4..8}, plain least squares
* ``ring256``       256 cameras, 60,000 points, tracks U{3..8}, Huber delta 1 px
* ``rings1500``     1,500 cameras on rings of radius 8/10/12, 500,000 points,
                    tracks U{3..10}, 2% outliers uniform over the image (only on points
                    that stay well triangulated by their inliers), Huber 1 px
* ``zipf4000``      4,000 cameras on one circle of radius 10, 1,000,000 points, Zipf(1.5)
                    tracks clipped to 2..200, focal U(400, 800), no outliers, Huber 1 px

Initial estimates perturb every rotation by N(0, 0.01 rad) per axis and every
translation and point coordinate by N(0, 0.05).
"""
from dataclasses import dataclass, field

import numpy as np


@dataclass
class BalProblem:
    name: str
    cameras: np.ndarray        # (nc, 9) float64, initial estimate
    points: np.ndarray         # (np, 3) float64, initial estimate
    obs_camera: np.ndarray     # (no,) int64
    obs_point: np.ndarray      # (no,) int64
    obs_pixel: np.ndarray      # (no, 2) float64
    huber: float = 0.0         # 0 = plain least squares
    truth_cameras: np.ndarray = field(default=None, repr=False)
    truth_points: np.ndarray = field(default=None, repr=False)

    @property
    def n_cameras(self):
        return self.cameras.shape[0]

    @property
    def n_points(self):
        return self.points.shape[0]

    @property
    def n_observations(self):
        return self.obs_camera.shape[0]

    def to_bal_text(self):
        """BAL text (read back bit for bit by fagra_amd.bal.parse)."""
        from . import bal

        return bal.format(self)


def rotation_matrices(w):
    """Rodrigues rotation for each row of w (n, 3)."""
    w = np.asarray(w, dtype=np.float64)
    theta = np.linalg.norm(w, axis=1)
    small = theta < 1e-12
    k = np.where(small[:, None], 0.0, w / np.where(small, 1.0, theta)[:, None])
    K = np.zeros((len(w), 3, 3))
    K[:, 0, 1], K[:, 0, 2] = -k[:, 2], k[:, 1]
    K[:, 1, 0], K[:, 1, 2] = k[:, 2], -k[:, 0]
    K[:, 2, 0], K[:, 2, 1] = -k[:, 1], k[:, 0]
    s, c = np.sin(theta)[:, None, None], np.cos(theta)[:, None, None]
    return np.eye(3)[None] + s * K + (1 - c) * (K @ K)


def angle_axis(R):
    """Inverse of rotation_matrices for rotations with angle < pi."""
    tr = np.trace(R, axis1=1, axis2=2)
    theta = np.arccos(np.clip((tr - 1) / 2, -1, 1))
    v = np.stack([R[:, 2, 1] - R[:, 1, 2], R[:, 0, 2] - R[:, 2, 0], R[:, 1, 0] - R[:, 0, 1]], axis=1)
    s = np.sin(theta)
    scale = np.where(s > 1e-12, theta / (2 * np.where(s > 1e-12, s, 1.0)), 0.5)
    return v * scale[:, None]


def project(cameras, points, obs_camera, obs_point):
    R = rotation_matrices(cameras[:, :3])
    P = np.einsum("nij,nj->ni", R[obs_camera], points[obs_point]) + cameras[obs_camera, 3:6]
    q = -P[:, :2] / P[:, 2:3]
    r2 = np.sum(q * q, axis=1)
    c = cameras[obs_camera]
    scale = c[:, 6] * (1 + c[:, 7] * r2 + c[:, 8] * r2 * r2)
    return scale[:, None] * q


def _ring_cameras(rng, n, radii, focal):
    per = np.array_split(np.arange(n), len(radii))
    centers = np.zeros((n, 3))
    for idx, radius in zip(per, radii):
        ang = 2 * np.pi * np.arange(len(idx)) / max(len(idx), 1) + rng.uniform(0, 2 * np.pi)
        centers[idx, 0] = radius * np.cos(ang)
        centers[idx, 1] = radius * np.sin(ang)
        centers[idx, 2] = rng.uniform(-1.0, 1.0, len(idx))
    z = centers / np.linalg.norm(centers, axis=1, keepdims=True)   # camera +z points away
    up = np.array([0.0, 0.0, 1.0])
    x = np.cross(up, z)
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    y = np.cross(z, x)
    R = np.stack([x, y, z], axis=1)                                 # rows: camera axes
    t = -np.einsum("nij,nj->ni", R, centers)
    cams = np.zeros((n, 9))
    cams[:, :3] = angle_axis(R)
    cams[:, 3:6] = t
    cams[:, 6] = focal
    cams[:, 7:9] = rng.normal(0.0, 1e-3, (n, 2))
    return cams


def _tracks(rng, n_points, n_cameras, lengths):
    """Distinct cameras per point, chosen uniformly without replacement."""
    offsets = np.concatenate([[0], np.cumsum(lengths)])
    total = int(offsets[-1])
    point = np.repeat(np.arange(n_points), lengths)
    cam = rng.integers(0, n_cameras, total)
    while True:
        order = np.lexsort((cam, point))
        sp, sc = point[order], cam[order]
        dup = np.zeros(total, bool)
        dup[1:] = (sp[1:] == sp[:-1]) & (sc[1:] == sc[:-1])
        if not dup.any():
            break
        bad = order[dup]
        cam[bad] = rng.integers(0, n_cameras, len(bad))
    return point, cam


def _triangulable(cams, pts, obs_cam, obs_point, inlier, n_points, min_inliers=3, min_angle_deg=30.0):
    """Points whose inlier observations pin them down: at least ``min_inliers`` of them, two
    of which see the point under a parallax angle in [min_angle, 180 - min_angle].

    Under the Huber cost an outlier pulls with a bounded force, and a point held only by two
    inlier rays of small parallax can lower its cost by sliding to infinity or behind a camera,
    where the inlier residuals stay bounded and the outlier's shrinks; outliers are only placed
    on points that this test keeps well triangulated. Observations are grouped by point.
    """
    R = rotation_matrices(cams[:, :3])
    centers = -np.einsum("nji,nj->ni", R, cams[:, 3:6])
    rays = pts[obs_point] - centers[obs_cam]
    rays /= np.linalg.norm(rays, axis=1, keepdims=True)
    idx = np.flatnonzero(inlier)
    pts_with, first = np.unique(obs_point[idx], return_index=True)
    ref = np.zeros((n_points, 3))
    ref[pts_with] = rays[idx[first]]
    cosang = np.abs(np.einsum("ni,ni->n", rays, ref[obs_point]))
    wide = inlier & (cosang <= np.cos(np.radians(min_angle_deg)))
    n_in = np.bincount(obs_point, weights=inlier, minlength=n_points)
    n_wide = np.bincount(obs_point, weights=wide, minlength=n_points)
    return (n_in >= min_inliers) & (n_wide >= 1)


CONFIGS = {
    "ring16": dict(n_cameras=16, n_points=2000, radii=(10.0,), track=("uniform", 4, 8), huber=0.0,
                   outliers=0.0, focal=("const", 500.0)),
    "ring256": dict(n_cameras=256, n_points=60000, radii=(10.0,), track=("uniform", 3, 8), huber=1.0,
                    outliers=0.0, focal=("const", 500.0)),
    "rings1500": dict(n_cameras=1500, n_points=500000, radii=(8.0, 10.0, 12.0), track=("uniform", 3, 10),
                      huber=1.0, outliers=0.02, focal=("const", 500.0)),
    "zipf4000": dict(n_cameras=4000, n_points=1000000, radii=(10.0,), track=("zipf", 1.5, 2, 200),
                     huber=1.0, outliers=0.0, focal=("uniform", 400.0, 800.0)),
}


def generate(name="ring16", seed=0, n_cameras=None, n_points=None, **overrides):
    """Generate a named workload; sizes and options can be overridden for tests."""
    cfg = dict(CONFIGS[name])
    cfg.update(overrides)
    nc = n_cameras or cfg["n_cameras"]
    npts = n_points or cfg["n_points"]
    rng = np.random.default_rng(seed)
    fk = cfg["focal"]
    focal = fk[1] if fk[0] == "const" else rng.uniform(fk[1], fk[2], nc)
    cams = _ring_cameras(rng, nc, cfg["radii"], focal)
    pts = rng.uniform(-2.0, 2.0, (npts, 3))
    tr = cfg["track"]
    if tr[0] == "uniform":
        lengths = rng.integers(tr[1], tr[2] + 1, npts)
    else:
        lengths = np.clip(rng.zipf(tr[1], npts), tr[2], tr[3])
    lengths = np.minimum(lengths, nc)
    obs_point, obs_cam = _tracks(rng, npts, nc, lengths)
    pixel = project(cams, pts, obs_cam, obs_point)
    pixel += rng.normal(0.0, 0.5, pixel.shape)
    if cfg["outliers"] > 0:
        # outliers go to observations of points that stay triangulable without them, at the
        # rate that puts the configured fraction over all observations
        everyone = np.ones(len(pixel), bool)
        eligible = (_triangulable(cams, pts, obs_cam, obs_point, everyone, npts) & (lengths >= 4))[obs_point]
        rate = min(1.0, cfg["outliers"] * len(pixel) / max(1, int(eligible.sum())))
        bad = eligible & (rng.random(len(pixel)) < rate)
        bad &= _triangulable(cams, pts, obs_cam, obs_point, ~bad, npts)[obs_point]
        half = np.abs(pixel).max()
        pixel[bad] = rng.uniform(-half, half, (int(bad.sum()), 2))
    init_c = cams.copy()
    init_c[:, :3] += rng.normal(0.0, 0.01, (nc, 3))
    init_c[:, 3:6] += rng.normal(0.0, 0.05, (nc, 3))
    init_p = pts + rng.normal(0.0, 0.05, pts.shape)
    return BalProblem(name=name, cameras=init_c, points=init_p, obs_camera=obs_cam.astype(np.int64),
                      obs_point=obs_point.astype(np.int64), obs_pixel=pixel, huber=float(cfg["huber"]),
                      truth_cameras=cams, truth_points=pts)


def parse_bal(text):
    """Problem::parse (examples/bal/model.rs:59-113) via the native reader."""
    from . import bal

    return bal.parse(text)
