"""The reference's on-disk formats: run configuration, tracer configuration,
OTF files, image lists / TIFF frames and object / track CSV files.

* ``BasicSetting.readConfig``  - BasicSetting::readConfig (Config.cpp:8-158)
* ``TracerConfig.readConfig``  - ObjectConfig::readCommonConfig +
  TracerConfig::readConfig (Config.cpp:161-276), including the OTF estimate
  from the first five frames (OTF::estimateUniformOTFFromImage,
  OTF.cpp:16-144), which runs on the GPU (peaks) through olpt_estimate_otf
* ``load_otf`` / ``save_otf``  - OTF::loadParam(file) / saveParam
  (OTF.cpp:195-271)
* ``ImageIO``                  - ImageIO::loadImgPath / loadImg
  (ImageIO.cpp:28-260): a text file of image paths, one TIFF per frame
* ``save_obj_info``            - IPR::saveObjInfo (IPR.cpp:307-343)

Parsing follows the reference's rules: lines are trimmed, a line starting
with '#' is skipped, anything after '#' is dropped, and numbers are read
with std::stoi / std::stod semantics (leading number of the field).
"""
from __future__ import annotations

import ctypes
import dataclasses
import os
import re

import numpy as np

from . import synth as S
from .camfile import load_camera

_INT = re.compile(r"\s*([+-]?\d+)")
_FLT = re.compile(r"\s*([+-]?(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?|[+-]?(?:inf|nan))", re.I)


def stoi(s: str) -> int:
    m = _INT.match(s)
    if not m:
        raise ValueError(f"stoi: no conversion: {s!r}")
    return int(m.group(1))


def stod(s: str) -> float:
    m = _FLT.match(s)
    if not m:
        raise ValueError(f"stod: no conversion: {s!r}")
    return float(m.group(1))


def config_lines(path: str) -> list[str]:
    """Non-empty, comment-stripped lines (Config.cpp:16-38)."""
    out = []
    with open(path) as f:
        for line in f:
            line = line.strip(" \t\r\n")
            if not line:
                continue
            p = line.find("#")
            if p == 0:
                continue
            if p >= 0:
                line = line[:p]
            line = line.strip(" \t\r\n")
            if line:
                out.append(line)
    return out


@dataclasses.dataclass
class BasicSetting:
    """BasicSetting (Config.h:14-37)."""
    n_cam: int = 0
    frame_start: int = 0
    frame_end: int = 0
    fps: int = 0
    n_thread: int = 0
    cam_list: list = dataclasses.field(default_factory=list)
    axis_limit: tuple = (0.0,) * 6
    voxel_to_mm: float = 0.0
    output_path: str = ""
    image_file_paths: list = dataclasses.field(default_factory=list)
    object_types: list = dataclasses.field(default_factory=list)
    object_config_paths: list = dataclasses.field(default_factory=list)
    load_track: bool = False
    load_track_frame: int = 0
    load_track_path: str = ""

    def readConfig(self, config_path: str) -> bool:  # noqa: N802 - reference name
        if not os.path.exists(config_path):
            return False
        lines = config_lines(config_path)
        if len(lines) < 11:
            return False
        it = iter(lines)
        a, b = next(it).split(",")[:2]
        self.frame_start, self.frame_end = stoi(a), stoi(b)
        self.fps = stoi(next(it))
        self.n_thread = stoi(next(it))
        self.n_cam = stoi(next(it))
        self.cam_list = []
        for _ in range(self.n_cam):
            fields = next(it).split(",")
            self.cam_list.append(load_camera(fields[0], float(stoi(fields[1]))))
        self.image_file_paths = [next(it) for _ in range(self.n_cam)]
        self.axis_limit = tuple(stod(v) for v in next(it).split(",")[:6])
        self.voxel_to_mm = stod(next(it))
        self.output_path = next(it)
        if self.output_path and self.output_path[-1] not in "/\\":
            self.output_path += "/"
        self.object_types = [t for t in next(it).split(",") if t]
        self.object_config_paths = [next(it) for _ in self.object_types]
        fields = next(it).split(",")
        self.load_track = stoi(fields[0]) != 0
        self.load_track_frame = self.frame_start
        if self.load_track:
            if len(fields) > 1 and fields[1]:
                self.load_track_frame = stoi(fields[1])
            rest = list(it)
            if rest:
                self.load_track_path = rest[0]
        if self.n_cam <= 0:
            raise ValueError(f"Invalid n_cam in config: {config_path}")
        if self.frame_end < self.frame_start:
            raise ValueError(f"Invalid frame range in config: {config_path}")
        if self.n_thread < 0:
            self.n_thread = 0
        return True


@dataclasses.dataclass
class TracerConfig:
    """ObjectConfig + TracerConfig parameters (Config.h:39-205), lengths in
    mm (the file gives them in voxels)."""
    radius_search_obj: float = 0.0
    n_initial_frames: int = 4
    radius_search_track: float = 0.0
    shake_width: float = 0.0
    shakewidth_min: float = 0.0
    pf_nx: int = 0
    pf_ny: int = 0
    pf_nz: int = 0
    pf_r: float = 0.0
    pf_nbin: tuple = (11, 11, 11)
    pf_smooth: bool = True
    pf_sigma: tuple = (1.0, 1.0, 1.0)
    n_loop_ipr: int = 4
    n_shake_loop: int = 4
    thred_ghost: float = 0.1
    ratio_augimg: int = 2
    tol_2d_px: float = 1.0
    tol_3d_mm: float = 2.4e-2
    match_cam_count: int = 4
    idmap_cell_px: int = 4
    n_cam_reduced: int = 1
    n_loop_ipr_reduced: int = 2
    n_obj2d_process_max: int = 200000
    min_obj_int: int = 0
    radius_obj: float = 2.0
    limit: tuple = (0.0,) * 6
    otf: "Otf | None" = None

    def readCommonConfig(self, path: str, settings: BasicSetting) -> tuple[list, int]:  # noqa: N802
        lines = config_lines(path)
        v = settings.voxel_to_mm
        it = 0

        def nxt():
            nonlocal it
            it += 1
            return lines[it - 1]
        try:
            self.radius_search_obj = stod(nxt()) * v
            self.n_initial_frames = max(4, stoi(nxt()))
            self.radius_search_track = stod(nxt()) * v
            self.shake_width = stod(nxt()) * v
            self.shakewidth_min = self.shake_width / 20
            self.limit = tuple(settings.axis_limit)
            self.pf_nx, self.pf_ny, self.pf_nz = stoi(nxt()), stoi(nxt()), stoi(nxt())
            self.pf_r = stod(nxt()) * v
            self.n_loop_ipr = stoi(nxt())
            self.n_shake_loop = stoi(nxt())
            self.thred_ghost = stod(nxt())
            self.tol_2d_px = stod(nxt())
            self.tol_3d_mm = stod(nxt()) * v
            self.match_cam_count = min(self.match_cam_count, settings.n_cam)
            self.n_cam_reduced = stoi(nxt())
            self.n_loop_ipr_reduced = stoi(nxt())
        except (IndexError, ValueError) as e:
            raise ValueError(f"Config Error: Failed to parse common config in {path}") from e
        return lines, it

    def readConfig(self, path: str, settings: BasicSetting, ctx=None,  # noqa: N802
                   n_img_otf: int = 5) -> bool:
        """Parameters, then the OTF: a 5x5x5 grid over the view volume with
        the default values (OTF::loadParam, OTF.cpp:159-193), each camera's
        nodes replaced by the estimate from the first n_img_otf frames when
        a context is given (Config.cpp:238-266)."""
        lines, it = self.readCommonConfig(path, settings)
        try:
            self.min_obj_int = stoi(lines[it])
            self.radius_obj = stod(lines[it + 1])
        except (IndexError, ValueError) as e:
            raise ValueError(f"Config Error: Failed to parse common config in {path}") from e
        self.otf = Otf.default(settings.n_cam, (5, 5, 5), settings.axis_limit,
                               [c.max_intensity for c in settings.cam_list])
        if ctx is not None:
            self.estimate_otf(ctx, settings, n_img_otf)
        return True

    def estimate_otf(self, ctx, settings: BasicSetting, n_img: int = 5) -> list[bool]:
        import torch
        from . import _lib
        ios = [ImageIO("", p) for p in settings.image_file_paths]
        frames = np.stack([np.stack([io.loadImg(settings.frame_start + j) for io in ios])
                           for j in range(n_img)])
        d, code = _device_frames(frames)
        abca = np.zeros((settings.n_cam, 4))
        ok = np.zeros(settings.n_cam, np.int32)
        p = self.ipr_params()
        torch.cuda.synchronize()
        _lib.check(_lib.lib().olpt_estimate_otf(
            ctx.h, ctypes_ptr(d), code, n_img, ctypes.byref(p),
            abca.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            ok.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        for k in range(settings.n_cam):
            if ok[k]:
                self.otf.set_uniform(k, abca[k])
        return [bool(v) for v in ok]

    @classmethod
    def from_ipr_params(cls, p: S.OlptIprParams, otf: "Otf | None" = None) -> "TracerConfig":
        """A configuration holding the IPR / stereo / shake fields of p (the
        inverse of ipr_params) and the given OTF."""
        cfg = cls()
        for f in ("n_cam_reduced", "n_loop_ipr", "n_loop_ipr_reduced", "n_obj2d_process_max",
                  "match_cam_count", "idmap_cell_px", "tol_2d_px", "tol_3d_mm", "shake_width",
                  "shakewidth_min", "n_shake_loop", "ratio_augimg", "thred_ghost", "min_obj_int",
                  "radius_obj"):
            setattr(cfg, f, getattr(p, f))
        cfg.limit = tuple(float(v) for v in p.limit)
        cfg.otf = otf
        return cfg

    def ipr_params(self) -> S.OlptIprParams:
        p = S.OlptIprParams()
        p.n_cam_reduced = self.n_cam_reduced
        p.n_loop_ipr = self.n_loop_ipr
        p.n_loop_ipr_reduced = self.n_loop_ipr_reduced
        p.n_obj2d_process_max = self.n_obj2d_process_max
        p.match_cam_count = self.match_cam_count
        p.idmap_cell_px = self.idmap_cell_px
        p.tol_2d_px = self.tol_2d_px
        p.tol_3d_mm = self.tol_3d_mm
        for i, v in enumerate(self.limit):
            p.limit[i] = v
        p.shake_width = self.shake_width
        p.shakewidth_min = self.shakewidth_min
        p.n_shake_loop = self.n_shake_loop
        p.ratio_augimg = self.ratio_augimg
        p.thred_ghost = self.thred_ghost
        p.min_obj_int = self.min_obj_int
        p.radius_obj = self.radius_obj
        return p

    def stb_params(self, settings: BasicSetting) -> S.OlptStbParams:
        q = S.OlptStbParams()
        q.frame_start, q.frame_end = settings.frame_start, settings.frame_end
        q.n_initial_frames = self.n_initial_frames
        q.radius_search_obj_mm = self.radius_search_obj
        q.radius_search_track_mm = self.radius_search_track
        for i, v in enumerate(settings.axis_limit):
            q.limit[i] = v
        q.pf_nx, q.pf_ny, q.pf_nz = self.pf_nx, self.pf_ny, self.pf_nz
        q.pf_r_mm = self.pf_r
        for i in range(3):
            q.pf_nbin[i] = self.pf_nbin[i]
            q.pf_sigma[i] = self.pf_sigma[i]
        q.pf_smooth = int(self.pf_smooth)
        return q


# ---------------------------------------------------------------------------
# OTF files
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class Otf:
    """OTFParam (OTF.h:16-31): per camera and grid node (a, b, c, alpha)."""
    n_cam: int
    shape: tuple
    boundary: tuple
    a: np.ndarray
    b: np.ndarray
    c: np.ndarray
    alpha: np.ndarray

    @classmethod
    def default(cls, n_cam, shape, boundary, max_intensity):
        """OTF::loadParam(n_cam, nx, ny, nz, ...) (OTF.cpp:159-193): a =
        max_intensity / 2, b = c = 1.5, alpha = 0; grids below 2 become 2."""
        shape = tuple(max(2, int(v)) for v in shape)
        ng = int(np.prod(shape))
        a = np.repeat(np.asarray(max_intensity, np.float64)[:, None] / 2.0, ng, axis=1)
        return cls(n_cam, shape, tuple(boundary), a, np.full((n_cam, ng), 1.5),
                   np.full((n_cam, ng), 1.5), np.zeros((n_cam, ng)))

    def set_uniform(self, cam: int, abca) -> None:
        self.a[cam, :], self.b[cam, :], self.c[cam, :], self.alpha[cam, :] = abca

    def to_c(self):
        """(olpt_otf struct, keepalive arrays)."""
        return S.make_otf_grid(self.n_cam, self.shape, self.boundary, self.a, self.b, self.c,
                               self.alpha)


def load_otf(path: str) -> Otf:
    """OTF::loadParam(file) (OTF.cpp:195-243): '#' at column 0 skips a line,
    later '#' truncates it; then n_cam, nx, ny, nz, n_grid, the boundary and
    the a, b, c, alpha matrices in reading order."""
    parts = []
    with open(path) as f:
        for line in f:
            line = line.rstrip("\n")
            p = line.find("#")
            if p == 0:
                continue
            if p > 0:
                line = line[:p]
            parts.append(line)
    nums = [float(t) for t in re.findall(r"[+-]?(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?",
                                         "\t".join(parts))]
    n_cam, nx, ny, nz, ng = (int(v) for v in nums[:5])
    bnd = tuple(nums[5:11])
    vals = np.asarray(nums[11:11 + 4 * n_cam * ng], np.float64)
    if vals.size != 4 * n_cam * ng:
        raise ValueError(f"{path}: expected {4 * n_cam * ng} OTF values, found {vals.size}")
    a, b, c, al = vals.reshape(4, n_cam, ng)
    return Otf(n_cam, (nx, ny, nz), bnd, a.copy(), b.copy(), c.copy(), al.copy())


def save_otf(otf: Otf, path: str) -> None:
    """OTF::saveParam (OTF.cpp:245-271); matrices as Matrix::write
    (scientific, 8 digits, Matrix.hpp:281-292)."""
    def mat(m):
        return "".join(",".join(f"{v:.8e}" for v in row) + "\n" for row in m)
    nx, ny, nz = otf.shape
    with open(path, "w") as f:
        f.write("# Size: (n_cam,nx,ny,nz,n_grid)\n")
        f.write(f"{otf.n_cam},{nx},{ny},{nz},{nx * ny * nz}\n")
        f.write("# Boundary: (xmin,xmax,ymin,ymax,zmin,zmax)\n")
        f.write(",".join(_g(v) for v in otf.boundary) + "\n")
        for name, m in (("a", otf.a), ("b", otf.b), ("c", otf.c), ("alpha", otf.alpha)):
            f.write(f"# {name}: (n_cam,n_grid)\n")
            f.write(mat(m))


def _g(v: float) -> str:
    """operator<< on a double with the default stream format (%g, 6 digits)."""
    s = f"{v:g}"
    return s


# ---------------------------------------------------------------------------
# images
# ---------------------------------------------------------------------------
class ImageIO:
    """ImageIO (ImageIO.cpp): a text file listing one image path per line
    (prefixed by folder_path); loadImg(i) reads line i's TIFF as its
    unsigned pixel type (8 or 16 bit, one channel, any strip/tile layout)."""

    def __init__(self, folder_path: str = "", file_img_path: str | None = None):
        self.img_path: list[str] = []
        if file_img_path is not None:
            self.loadImgPath(folder_path, file_img_path)

    def loadImgPath(self, folder_path: str, file_img_path: str) -> None:  # noqa: N802
        with open(folder_path + file_img_path) as f:
            self.img_path = [folder_path + line.rstrip("\n") for line in f]

    def loadImg(self, img_id: int) -> np.ndarray:  # noqa: N802
        if img_id >= len(self.img_path):
            raise IndexError(f"Image id: {img_id} is larger than total number of image: "
                             f"{len(self.img_path)}")
        return read_tiff(self.img_path[img_id])


def read_tiff(path: str) -> np.ndarray:
    from PIL import Image
    with Image.open(path) as im:
        if im.mode == "L":
            return np.asarray(im, dtype=np.uint8).copy()
        if im.mode in ("I;16", "I;16B", "I;16L"):
            return np.asarray(im, dtype=np.uint16).copy()
        raise ValueError(f"{path}: unsupported TIFF mode {im.mode} (8/16-bit grayscale only)")


def write_tiff(path: str, img: np.ndarray) -> None:
    """Uncompressed single-strip TIFF of an 8- or 16-bit image."""
    from PIL import Image
    if img.dtype not in (np.uint8, np.uint16):
        raise ValueError("write_tiff takes uint8 or uint16 images")
    im = Image.fromarray(np.ascontiguousarray(img.astype(img.dtype.newbyteorder("<"))))
    if im.mode not in ("L", "I;16"):
        raise ValueError(f"unexpected TIFF mode {im.mode}")
    im.save(path, format="TIFF")


def _device_frames(frames: np.ndarray):
    from .context import to_device_image
    return to_device_image(frames)


def ctypes_ptr(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


# ---------------------------------------------------------------------------
# object / track files
# ---------------------------------------------------------------------------
def save_obj_info(filename: str, xyz: np.ndarray, uv: np.ndarray) -> None:
    """IPR::saveObjInfo (IPR.cpp:307-343) for tracers: header with
    Error,Ncam columns that the rows do not fill (as the reference writes
    it), rows X,Y,Z then (x, y) per camera in the default stream format
    (Tracer3D::saveObject3D, ObjectInfo.cpp:76-94); a camera without an
    observation (NaN x) keeps its two columns empty."""
    xyz = np.asarray(xyz, np.float64).reshape(-1, 3)
    n_cam = uv.shape[1] if uv.ndim == 3 else 0
    with open(filename, "w") as f:
        f.write("WorldX,WorldY,WorldZ,Error,Ncam")
        for i in range(n_cam):
            f.write(f",cam{i}_x(col),cam{i}_y(row)")
        f.write("\n")
        for j in range(len(xyz)):
            row = [_g(v) for v in xyz[j]]
            for i in range(n_cam):
                row += (["", ""] if np.isnan(uv[j, i, 0])
                        else [_g(uv[j, i, 0]), _g(uv[j, i, 1])])
            f.write(",".join(row) + "\n")


def write_basic_config(path: str, *, frame_range, fps, n_thread, cam_files, max_intensity,
                       image_lists, axis_limit, voxel_to_mm, output_path, object_types,
                       object_configs, load_track=(0, -1), load_track_path="") -> None:
    """A run configuration in the layout BasicSetting::readConfig reads
    (Config.cpp:44-129), with the reference sample's comment lines."""
    L = ["# Frame Range: [startID,endID]", f"{frame_range[0]},{frame_range[1]}",
         "# Frame Rate: [Hz]", str(fps), "# Number of Threads: (0: use as many as possible)",
         str(n_thread), "# Number of Cameras: ", str(len(cam_files)),
         "# Camera File Path, Max Intensity"]
    L += [f"{c},{int(m)}" for c, m in zip(cam_files, max_intensity)]
    L += ["# Image File Path"] + list(image_lists)
    L += ["# View Volume: (xmin,xmax,ymin,ymax,zmin,zmax)", ",".join(repr(float(v)) for v in axis_limit),
          "# Voxel to MM: e.g. use 1000^3 voxel, (xmax-xmin)/1000", repr(float(voxel_to_mm)),
          "# Output Folder Path: ", output_path, "# Object Types: ", ",".join(object_types),
          "# STB Config Files:"] + list(object_configs)
    L += ["# Flag to load previous track files, previous frameID",
          f"{int(load_track[0])},{int(load_track[1])}"]
    if load_track_path:
        L += ["# Path to track files", load_track_path]
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def write_tracer_config(path: str, *, radius_search_obj_vox, n_initial_frames,
                        radius_search_track_vox, shake_width_vox, pf_grid, pf_r_vox,
                        n_loop_ipr, n_shake_loop, thred_ghost, tol_2d_px, tol_3d_vox,
                        n_cam_reduced, n_loop_ipr_reduced, min_obj_int, radius_obj) -> None:
    """A tracer configuration in the layout of readCommonConfig +
    TracerConfig::readConfig (Config.cpp:191-244), as the reference sample
    tracerConfig.txt lays it out."""
    L = ["######### Initial Phase #########", f"{radius_search_obj_vox} # Search radius for connecting tracks to objects",
         f"{n_initial_frames} # Number of frames for initial phase",
         "######### Convergence Phase #########", f"{radius_search_track_vox} # Avg Interparticle spacing. (vox)",
         "######### Shake #########", f"{shake_width_vox} # shake width",
         "######### Predict Field #########", f"{pf_grid[0]} # xgrid", f"{pf_grid[1]} # ygrid",
         f"{pf_grid[2]} # zgrid", f"{pf_r_vox} # searchRadius [voxel]",
         "######### IPR #########", f"{n_loop_ipr} # No. of IPR loop", f"{n_shake_loop} # No. of Shake loop",
         f"{thred_ghost} # ghost threshold", f"{tol_2d_px} # 2D tolerance [px]",
         f"{tol_3d_vox} # 3D tolerance [voxel]", f"{n_cam_reduced} # number of reduced camera",
         f"{n_loop_ipr_reduced} # no. of ipr loops for each reduced camera combination",
         "######### Object Info #########", f"{min_obj_int} # 2D particle finder threshold",
         f"{radius_obj} # Particle radius [px]"]
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def write_run_folder(folder: str, cams, frames: np.ndarray, *, axis_limit, voxel_to_mm,
                     tracer: dict, fps: int = 1, n_thread: int = 0,
                     n_tiff: int | None = None) -> str:
    """Everything a run reads, under folder: camera files, one TIFF per
    camera and frame with its image list, tracerConfig.txt and config.txt
    (output folder: folder/out/).  frames: [F][n_cam][H][W] uint8/uint16.
    n_tiff: write only the first n_tiff frames' TIFFs (the lists still name
    every frame), for runs that take later frames from memory.  Returns the
    path of config.txt (all paths in it are absolute)."""
    from .camfile import save_camera
    folder = os.path.abspath(folder)
    os.makedirs(os.path.join(folder, "camFile"), exist_ok=True)
    cam_files, lists = [], []
    for k, cam in enumerate(cams):
        cf = os.path.join(folder, "camFile", f"cam{k + 1}.txt")
        save_camera(cam, cf)
        cam_files.append(cf)
        d = os.path.join(folder, "imgFile", f"cam{k + 1}")
        os.makedirs(d, exist_ok=True)
        names = []
        for f in range(frames.shape[0]):
            pth = os.path.join(d, f"img{f:05d}.tif")
            if n_tiff is None or f < n_tiff:
                write_tiff(pth, frames[f, k])
            names.append(pth)
        lf = os.path.join(folder, "imgFile", f"cam{k + 1}ImageNames.txt")
        with open(lf, "w") as fh:
            fh.write("\n".join(names) + "\n")
        lists.append(lf)
    tc = os.path.join(folder, "tracerConfig.txt")
    write_tracer_config(tc, **tracer)
    cfg = os.path.join(folder, "config.txt")
    write_basic_config(cfg, frame_range=(0, frames.shape[0] - 1), fps=fps, n_thread=n_thread,
                       cam_files=cam_files, max_intensity=[c.max_intensity for c in cams],
                       image_lists=lists, axis_limit=axis_limit, voxel_to_mm=voxel_to_mm,
                       output_path=os.path.join(folder, "out") + "/", object_types=["Tracer"],
                       object_configs=[tc])
    return cfg


def read_tracks(path: str) -> dict:
    """A track file written by STB::saveTracks (STB.cpp:967-1004): returns
    {'track': int32, 'frame': int32, 'xyz': float64 [n][3], 'uv': [n][n_cam][2]}."""
    with open(path) as f:
        header = f.readline()
        n_cam = header.count("_x(col)")
        rows = [line.rstrip("\n").split(",") for line in f if line.strip()]
    if not rows:
        return dict(track=np.zeros(0, np.int32), frame=np.zeros(0, np.int32),
                    xyz=np.zeros((0, 3)), uv=np.zeros((0, n_cam, 2)))
    track = np.array([int(r[0]) for r in rows], np.int32)
    frame = np.array([int(r[1]) for r in rows], np.int32)
    vals = np.array([[float(v) for v in r[2:5 + 2 * n_cam]] for r in rows])
    return dict(track=track, frame=frame, xyz=vals[:, :3], uv=vals[:, 3:].reshape(-1, n_cam, 2))

