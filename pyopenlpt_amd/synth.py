"""Seeded synthetic IPR workloads: camera rig, particle fields, rendering.

The rig and rendering follow the reference's own fixture generator
(test/inputs/test_IPR/getTiffImg.py:9-45): each particle is projected with
the pinhole + OpenCV-distortion model, rendered with the Gaussian OTF
I = a*exp(-(b*xx^2 + c*yy^2)) on the 5x5 window [floor(u-2), floor(u+3)),
fused by per-pixel MAXIMUM, then (unlike the fixture) Gaussian noise is added
before clipping to [0, max] and truncation to the integer pixel type.

Cameras are returned as numpy records with the layout of ``olpt_camera``
(include/olpt_types.h) so they can be handed to every C entry point.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math

import numpy as np

# ---------------------------------------------------------------------------
# ctypes mirrors of include/olpt_types.h
# ---------------------------------------------------------------------------


class OlptCamera(ctypes.Structure):
    _fields_ = [
        ("n_row", ctypes.c_int32), ("n_col", ctypes.c_int32),
        ("n_dist", ctypes.c_int32), ("is_distorted", ctypes.c_int32),
        ("fx", ctypes.c_double), ("fy", ctypes.c_double),
        ("cx", ctypes.c_double), ("cy", ctypes.c_double),
        ("dist", ctypes.c_double * 12),
        ("R", ctypes.c_double * 9), ("T", ctypes.c_double * 3),
        ("Rinv", ctypes.c_double * 9), ("Tinv", ctypes.c_double * 3),
        ("max_intensity", ctypes.c_double),
    ]


class OlptOtf(ctypes.Structure):
    _fields_ = [
        ("n_cam", ctypes.c_int32), ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("boundary", ctypes.c_double * 6),
        ("a", ctypes.POINTER(ctypes.c_double)),
        ("b", ctypes.POINTER(ctypes.c_double)),
        ("c", ctypes.POINTER(ctypes.c_double)),
        ("alpha", ctypes.POINTER(ctypes.c_double)),
    ]


class OlptIprParams(ctypes.Structure):
    _fields_ = [
        ("n_cam_reduced", ctypes.c_int32), ("n_loop_ipr", ctypes.c_int32),
        ("n_loop_ipr_reduced", ctypes.c_int32),
        ("n_obj2d_process_max", ctypes.c_int32),
        ("match_cam_count", ctypes.c_int32), ("idmap_cell_px", ctypes.c_int32),
        ("tol_2d_px", ctypes.c_double), ("tol_3d_mm", ctypes.c_double),
        ("limit", ctypes.c_double * 6),
        ("shake_width", ctypes.c_double), ("shakewidth_min", ctypes.c_double),
        ("n_shake_loop", ctypes.c_int32), ("ratio_augimg", ctypes.c_int32),
        ("thred_ghost", ctypes.c_double),
        ("min_obj_int", ctypes.c_int32), ("reserved0", ctypes.c_int32),
        ("radius_obj", ctypes.c_double),
    ]


# ---------------------------------------------------------------------------
# Camera model (numpy, FP64) - same equations as Camera.cpp:680-689, :913-966
# ---------------------------------------------------------------------------


class OlptStbParams(ctypes.Structure):
    """olpt_stb_params (include/olpt_c.h): STBParam + PFParam + frame range."""
    _fields_ = [
        ("frame_start", ctypes.c_int32), ("frame_end", ctypes.c_int32),
        ("n_initial_frames", ctypes.c_int32),
        ("radius_search_obj_mm", ctypes.c_double),
        ("radius_search_track_mm", ctypes.c_double),
        ("limit", ctypes.c_double * 6),
        ("pf_nx", ctypes.c_int32), ("pf_ny", ctypes.c_int32), ("pf_nz", ctypes.c_int32),
        ("pf_r_mm", ctypes.c_double),
        ("pf_nbin", ctypes.c_int32 * 3),
        ("pf_smooth", ctypes.c_int32),
        ("pf_sigma", ctypes.c_double * 3),
    ]


class OlptTuning(ctypes.Structure):
    """olpt_tuning (include/olpt_c.h): result-neutral launch shapes."""
    _fields_ = [("n_streams", ctypes.c_int32), ("shake_blocks", ctypes.c_int32),
                ("build_blocks", ctypes.c_int32), ("prune_blocks", ctypes.c_int32),
                ("frames_per_batch", ctypes.c_int32)]

    def as_dict(self) -> dict:
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


@dataclasses.dataclass
class PinholeCamera:
    n_row: int
    n_col: int
    fx: float
    fy: float
    cx: float
    cy: float
    dist: np.ndarray          # 12 coefficients (unused tail = 0)
    n_dist: int
    R: np.ndarray             # 3x3
    T: np.ndarray             # 3
    Rinv: np.ndarray
    Tinv: np.ndarray
    max_intensity: float = 255.0

    @property
    def is_distorted(self) -> bool:
        # Camera.cpp:61-64: distortion is on iff a coefficient is > 1e-8
        return bool(np.any(self.dist[: self.n_dist] > 1e-8))

    def to_c(self) -> OlptCamera:
        c = OlptCamera()
        c.n_row, c.n_col = self.n_row, self.n_col
        c.n_dist = self.n_dist
        c.is_distorted = int(self.is_distorted)
        c.fx, c.fy, c.cx, c.cy = self.fx, self.fy, self.cx, self.cy
        for i in range(12):
            c.dist[i] = float(self.dist[i])
        for i in range(9):
            c.R[i] = float(self.R.flat[i])
            c.Rinv[i] = float(self.Rinv.flat[i])
        for i in range(3):
            c.T[i] = float(self.T[i])
            c.Tinv[i] = float(self.Tinv[i])
        c.max_intensity = float(self.max_intensity)
        return c

    def project(self, xyz: np.ndarray) -> np.ndarray:
        """World (N,3) mm -> pixel (N,2) (x=col, y=row)."""
        p = xyz @ self.R.T + self.T
        z = p[:, 2]
        x = np.where(z != 0, p[:, 0] / np.where(z != 0, z, 1), p[:, 0])
        y = np.where(z != 0, p[:, 1] / np.where(z != 0, z, 1), p[:, 1])
        if self.is_distorted:
            k = self.dist
            r2 = x * x + y * y
            r4 = r2 * r2
            r6 = r4 * r2
            a1 = 2 * x * y
            a2 = r2 + 2 * x * x
            a3 = r2 + 2 * y * y
            cdist = 1 + k[0] * r2 + k[1] * r4
            if self.n_dist > 4:
                cdist = cdist + k[4] * r6
            icd = 1.0
            if self.n_dist > 5:
                icd = 1.0 / (1 + k[5] * r2 + k[6] * r4 + k[7] * r6)
            xd = x * cdist * icd + k[2] * a1 + k[3] * a2
            yd = y * cdist * icd + k[2] * a3 + k[3] * a1
            if self.n_dist > 8:
                xd = xd + k[8] * r2 + k[9] * r4
                yd = yd + k[10] * r2 + k[11] * r4
            x, y = xd, yd
        return np.stack([x * self.fx + self.cx, y * self.fy + self.cy], axis=1)


def _look_at(center: np.ndarray) -> np.ndarray:
    """Rotation whose rows are the camera axes (x right, y down, z forward)
    for a camera at `center` looking at the origin."""
    z = -center / np.linalg.norm(center)
    up = np.array([0.0, 1.0, 0.0])
    x = np.cross(z, up)
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    return np.stack([x, y, z])


def make_rig(n_row: int = 1024, n_col: int = 1024, n_cam: int = 4,
             distance: float = 400.0, half: float = 20.0, fill: float = 0.9,
             dist_coeff=(0.0, 0.0, 0.0, 0.0, 0.0),
             max_intensity: float = 255.0) -> list[PinholeCamera]:
    """4 cameras at (az, el) = (+-30 deg, +-30 deg) around the cube
    [-half, half]^3, `distance` mm from its centre, focal length chosen so the
    projected cube spans `fill` of the image."""
    angles = [(30, 30), (-30, 30), (30, -30), (-30, -30), (0, 45), (0, -45),
              (45, 0), (-45, 0)][:n_cam]
    corners = np.array([[sx, sy, sz] for sx in (-half, half)
                        for sy in (-half, half) for sz in (-half, half)])
    cams = []
    for az, el in angles:
        a, e = math.radians(az), math.radians(el)
        center = distance * np.array([math.cos(e) * math.sin(a), math.sin(e),
                                      math.cos(e) * math.cos(a)])
        R = _look_at(center)
        T = -R @ center
        pc = corners @ R.T + T
        ext = np.max(np.abs(pc[:, :2] / pc[:, 2:3]))
        f = fill * 0.5 * min(n_row, n_col) / ext
        dist = np.zeros(12)
        dist[: len(dist_coeff)] = dist_coeff
        Rinv = R.T.copy()
        cams.append(PinholeCamera(
            n_row=n_row, n_col=n_col, fx=f, fy=f, cx=(n_col - 1) / 2.0,
            cy=(n_row - 1) / 2.0, dist=dist, n_dist=len(dist_coeff), R=R,
            T=T, Rinv=Rinv, Tinv=-Rinv @ T, max_intensity=max_intensity))
    return cams


# ---------------------------------------------------------------------------
# Particle fields
# ---------------------------------------------------------------------------


def uniform_particles(n: int, seed: int, half: float = 20.0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.uniform(-half, half, size=(n, 3))


def clustered_particles(n: int, seed: int, n_clusters: int = 8,
                        sigma: float = 5.0, half: float = 20.0,
                        centre_half: float = 15.0) -> np.ndarray:
    """Mixture of isotropic Gaussians (centres uniform in
    [-centre_half, centre_half]^3), clipped to the cube."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-centre_half, centre_half, size=(n_clusters, 3))
    lab = rng.integers(0, n_clusters, size=n)
    pts = centres[lab] + rng.normal(0.0, sigma, size=(n, 3))
    return np.clip(pts, -half, half)


# ---------------------------------------------------------------------------
# Rendering
# ---------------------------------------------------------------------------


def render(cams: list[PinholeCamera], xyz: np.ndarray, a, b: float, c: float,
           alpha: float = 0.0, noise_sigma: float = 0.0, seed: int = 0,
           dtype=np.uint8, radius: int = 2) -> np.ndarray:
    """Max-fusion Gaussian rendering -> (n_cam, H, W) integer images.

    `a` may be a scalar or a per-particle array."""
    n = xyz.shape[0]
    amp = np.broadcast_to(np.asarray(a, dtype=np.float64), (n,))
    rng = np.random.default_rng(seed)
    H, W = cams[0].n_row, cams[0].n_col
    out = np.empty((len(cams), H, W), dtype=dtype)
    vmax = float(np.iinfo(dtype).max) if np.issubdtype(dtype, np.integer) else np.inf
    cs, sn = math.cos(alpha), math.sin(alpha)
    offs = np.arange(2 * radius + 1)
    for k, cam in enumerate(cams):
        uv = cam.project(xyz)
        img = np.zeros((H, W), dtype=np.float64)
        x0 = np.floor(uv[:, 0] - radius).astype(np.int64)
        y0 = np.floor(uv[:, 1] - radius).astype(np.int64)
        cols = x0[:, None, None] + offs[None, None, :]
        rows = y0[:, None, None] + offs[None, :, None]
        cols, rows = np.broadcast_arrays(cols, rows)
        dx = cols - uv[:, 0, None, None]
        dy = rows - uv[:, 1, None, None]
        xx = dx * cs + dy * sn
        yy = -dx * sn + dy * cs
        val = amp[:, None, None] * np.exp(-(b * xx * xx + c * yy * yy))
        ok = (cols >= 0) & (cols < W) & (rows >= 0) & (rows < H)
        np.maximum.at(img, (rows[ok], cols[ok]), val[ok])
        if noise_sigma > 0:
            img += rng.normal(0.0, noise_sigma, size=img.shape)
        np.clip(img, 0.0, min(vmax, cam.max_intensity), out=img)
        out[k] = img.astype(dtype)
    return out


# ---------------------------------------------------------------------------
# Workload presets
# ---------------------------------------------------------------------------


@dataclasses.dataclass
class Workload:
    name: str
    cams: list
    images: np.ndarray          # (n_cam, H, W) integer
    truth: np.ndarray           # (N, 3)
    params: OlptIprParams
    otf_abca: tuple             # (a, b, c, alpha) uniform OTF values
    otf_boundary: tuple
    voxel_mm: float


def default_params(voxel_mm: float, *, tol_2d_px=1.0, n_loop_ipr=4,
                   n_shake_loop=4, min_obj_int=30, radius_obj=2.0,
                   n_cam_reduced=0, n_loop_ipr_reduced=2, half=20.0,
                   match_cam_count=4) -> OlptIprParams:
    """Tracer parameters from the reference sample tracerConfig.txt scaled by
    the voxel size: shake width 0.25 vox (shakewidth_min = width/20,
    Config.cpp:209), tol_3d 1 vox, ghost 0.1, ratio_augimg 2, cell 4 px."""
    p = OlptIprParams()
    p.n_cam_reduced = n_cam_reduced
    p.n_loop_ipr = n_loop_ipr
    p.n_loop_ipr_reduced = n_loop_ipr_reduced
    p.n_obj2d_process_max = 200000
    p.match_cam_count = match_cam_count
    p.idmap_cell_px = 4
    p.tol_2d_px = tol_2d_px
    p.tol_3d_mm = 1.0 * voxel_mm
    for i, v in enumerate((-half, half, -half, half, -half, half)):
        p.limit[i] = v
    p.shake_width = 0.25 * voxel_mm
    p.shakewidth_min = p.shake_width / 20.0
    p.n_shake_loop = n_shake_loop
    p.ratio_augimg = 2
    p.thred_ghost = 0.1
    p.min_obj_int = min_obj_int
    p.radius_obj = radius_obj
    return p


def make_otf(n_cam: int, abca=(125.0, 1.5, 1.5, 0.0), half: float = 20.0,
             n: int = 2):
    """Uniform OTF on an n^3 grid.  Returns (struct, keepalive arrays)."""
    ng = n * n * n
    arrs = [np.full(n_cam * ng, float(v)) for v in abca]
    o = OlptOtf()
    o.n_cam, o.nx, o.ny, o.nz = n_cam, n, n, n
    for i, v in enumerate((-half, half, -half, half, -half, half)):
        o.boundary[i] = v
    o.a, o.b, o.c, o.alpha = [x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
                              for x in arrs]
    return o, arrs


def make_otf_grid(n_cam: int, shape: tuple, boundary, a, b, c, alpha):
    """OTF on an nx x ny x nz grid from per-node arrays [n_cam][nx*ny*nz]
    (node index ix*ny*nz + iy*nz + iz).  Returns (struct, keepalive arrays)."""
    nx, ny, nz = shape
    ng = nx * ny * nz
    arrs = [np.ascontiguousarray(np.asarray(v, np.float64).reshape(n_cam * ng)) for v in (a, b, c, alpha)]
    o = OlptOtf()
    o.n_cam, o.nx, o.ny, o.nz = n_cam, nx, ny, nz
    for i, v in enumerate(boundary):
        o.boundary[i] = v
    o.a, o.b, o.c, o.alpha = [x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for x in arrs]
    return o, arrs


def workload_c1(seed: int = 1234, n: int = 2000, size: int = 512) -> Workload:
    """CPU-parity case: 512^2 u8, 2000 uniform particles, a=125, b=c=1.5,
    noise sigma 2, k1 = 1e-8 (loads as undistorted), 4 loops x 4 shakes."""
    cams = make_rig(size, size, dist_coeff=(1e-8, 0, 0, 0, 0), max_intensity=255)
    pts = uniform_particles(n, seed)
    img = render(cams, pts, 125.0, 1.5, 1.5, 0.0, 2.0, seed + 1, np.uint8)
    vox = 40.0 / (0.95 * size)
    p = default_params(vox, tol_2d_px=1.0, n_loop_ipr=4, n_shake_loop=4,
                       min_obj_int=30)
    return Workload("c1_512_u8_2000", cams, img, pts, p, (125.0, 1.5, 1.5, 0.0),
                    (-20, 20) * 3, vox)


def workload_c2(seed: int = 1234, n: int = 20000, size: int = 1024) -> Workload:
    """1024^2 u16, 20000 uniform particles, a~U(800,1200), b=c=1.2,
    noise sigma 10, 4 loops x 6 shakes, OTF nominal a=1000."""
    cams = make_rig(size, size, max_intensity=65535)
    pts = uniform_particles(n, seed)
    rng = np.random.default_rng(seed + 7)
    amp = rng.uniform(800.0, 1200.0, size=n)
    img = render(cams, pts, amp, 1.2, 1.2, 0.0, 10.0, seed + 1, np.uint16)
    vox = 40.0 / (0.95 * size)
    p = default_params(vox, tol_2d_px=1.0, n_loop_ipr=4, n_shake_loop=6,
                       min_obj_int=240)
    return Workload("c2_1024_u16_20000", cams, img, pts, p,
                    (1000.0, 1.2, 1.2, 0.0), (-20, 20) * 3, vox)


def workload_c3_frame(frame: int, seed: int = 1234, ppp: float = 0.05,
                      size: int = 1024) -> Workload:
    """One high-density frame: ppp 0.05 (52429 particles at 1024^2) from 8
    Gaussian clusters, otherwise as workload_c2."""
    n = int(round(ppp * size * size))
    cams = make_rig(size, size, max_intensity=65535)
    pts = clustered_particles(n, seed + frame)
    rng = np.random.default_rng(seed + frame + 7)
    amp = rng.uniform(800.0, 1200.0, size=n)
    img = render(cams, pts, amp, 1.2, 1.2, 0.0, 10.0, seed + frame + 1,
                 np.uint16)
    vox = 40.0 / (0.95 * size)
    p = default_params(vox, tol_2d_px=1.0, n_loop_ipr=4, n_shake_loop=6,
                       min_obj_int=240)
    return Workload(f"c3_1024_u16_ppp{ppp}_f{frame}", cams, img, pts, p,
                    (1000.0, 1.2, 1.2, 0.0), (-20, 20) * 3, vox)


def workload_noise_limit(seed: int = 4242, size: int = 512, cap: int = 3000) -> Workload:
    """3,000 particles (a = 125, b = c = 1.5) rendered on 4 x 512^2 u8 under
    uniform integer noise in [0, 70]: about one pixel in five is a local
    maximum, so each camera has ~49,000 peaks, above the 46,340 up to which
    std::shuffle draws two positions at a time; limit2DObjectNumber
    (n_obj2d_process_max = cap) keeps `cap` of them, and which ones decides
    what is reconstructed."""
    cams = make_rig(size, size, max_intensity=255)
    pts = uniform_particles(3000, seed)
    img = render(cams, pts, 125.0, 1.5, 1.5, 0.0, 2.0, seed + 1, np.uint8)
    noise = np.random.default_rng(seed + 2).integers(0, 71, size=img.shape)
    img = np.clip(img.astype(np.int32) + noise, 0, 255).astype(np.uint8)
    vox = 40.0 / (0.95 * size)
    p = default_params(vox, min_obj_int=30)
    p.n_obj2d_process_max = cap
    return Workload(f"noise_{size}_u8_limit{cap}", cams, img, pts, p,
                    (125.0, 1.5, 1.5, 0.0), (-20, 20) * 3, vox)


# ---------------------------------------------------------------------------
# Tracking sequences: particles advected by an ABC flow
# ---------------------------------------------------------------------------
def abc_velocity(p: np.ndarray, A: float = 1.0, B: float = 1.0, C: float = 1.0,
                 period_mm: float = 40.0) -> np.ndarray:
    """Arnold-Beltrami-Childress flow with period 2*pi <-> period_mm."""
    k = 2.0 * math.pi / period_mm
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    return np.stack([A * np.sin(k * z) + C * np.cos(k * y),
                     B * np.sin(k * x) + A * np.cos(k * z),
                     C * np.sin(k * y) + B * np.cos(k * x)], axis=1)


def abc_sequence(n: int, n_frames: int, seed: int, max_disp_mm: float,
                 half: float = 20.0) -> list[tuple[np.ndarray, np.ndarray]]:
    """Positions [n][3] and particle ids per frame: RK4 steps of the ABC
    flow scaled so that no particle moves more than max_disp_mm per frame
    (|u| <= 2 sqrt(3)); a particle leaving the cube is replaced by a new one
    (new id) at a uniform random position."""
    rng = np.random.default_rng(seed)
    p = rng.uniform(-half, half, (n, 3))
    ids = np.arange(n)
    next_id = n
    dt = max_disp_mm / (2.0 * math.sqrt(3.0))
    out = []
    for _ in range(n_frames):
        out.append((p.copy(), ids.copy()))
        k1 = abc_velocity(p)
        k2 = abc_velocity(p + 0.5 * dt * k1)
        k3 = abc_velocity(p + 0.5 * dt * k2)
        k4 = abc_velocity(p + dt * k3)
        p = p + dt / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)
        gone = np.any(np.abs(p) > half, axis=1)
        m = int(gone.sum())
        if m:
            p[gone] = rng.uniform(-half, half, (m, 3))
            ids[gone] = np.arange(next_id, next_id + m)
            next_id += m
    return out


def render_sequence(cams, seq, amp_range=(105.0, 145.0), b=1.5, c=1.5, noise=2.0,
                    seed=0, dtype=np.uint8) -> np.ndarray:
    """Frames [F][n_cam][H][W]: each particle keeps its amplitude (drawn per
    id) across frames."""
    rng = np.random.default_rng(seed)
    max_id = max(int(ids.max()) for _, ids in seq) + 1
    amp = rng.uniform(amp_range[0], amp_range[1], max_id)
    return np.stack([render(cams, p, amp[ids], b, c, 0.0, noise, seed + 1 + f, dtype)
                     for f, (p, ids) in enumerate(seq)])


# ---------------------------------------------------------------------------
# Quality against ground truth
# ---------------------------------------------------------------------------
def ghost_eval(xyz: np.ndarray, truth: np.ndarray, tol_mm: float) -> dict:
    """Reconstruction quality with find_ghosts semantics
    (demo/openLPT_utils/ghost_eval.py:72-92): each reconstruction's nearest
    ground-truth particle; within tol_mm (inclusive) it is real, otherwise a
    ghost.  recon_rate = distinct particles matched / N_truth, ghost_rate =
    ghosts / N_rec, rmse_mm over the real reconstructions."""
    from scipy.spatial import cKDTree
    if len(xyz) == 0:
        return dict(rmse_mm=None, recon_rate=0.0, ghost_rate=None, n_rec=0)
    d, i = cKDTree(truth).query(np.asarray(xyz, np.float64))
    real = d <= tol_mm
    return dict(rmse_mm=float(np.sqrt(np.mean(d[real] ** 2))) if real.any() else None,
                recon_rate=float(len(np.unique(i[real])) / len(truth)),
                ghost_rate=float(1.0 - real.mean()), n_rec=int(len(xyz)))
