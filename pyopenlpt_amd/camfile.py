"""Reference camera files (PINHOLE model) - Camera::loadParameters
(Camera.cpp:19-99, :312-344) and the inverse writer.

Comment text after '#' is dropped, the rest is read as a token stream:
type, two calibration-error tokens, "n_row,n_col", the 3x3 camera matrix,
the distortion list, the rotation vector, R (3x3), R^-1 (3x3), T, T^-1.
Numbers inside a token are comma-separated.
"""
from __future__ import annotations

import re

import numpy as np

from .synth import PinholeCamera

_NUM = re.compile(r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?")


class _Stream:
    def __init__(self, text: str):
        self.s = text
        self.i = 0

    def _skip_ws(self):
        while self.i < len(self.s) and self.s[self.i].isspace():
            self.i += 1

    def word(self) -> str:
        self._skip_ws()
        j = self.i
        while j < len(self.s) and not self.s[j].isspace():
            j += 1
        w = self.s[self.i:j]
        self.i = j
        return w

    def number(self) -> float:
        """`is >> double` followed by skipping one ',' (Matrix.hpp:102-116)."""
        self._skip_ws()
        m = _NUM.match(self.s, self.i)
        if not m:
            raise ValueError(f"expected a number at offset {self.i}")
        self.i = m.end()
        if self.i < len(self.s) and self.s[self.i] == ",":
            self.i += 1
        return float(m.group(0))


def _strip_comments(text: str) -> str:
    out = []
    for line in text.splitlines():
        pos = line.find("#")
        if pos == 0:
            continue
        if pos > 0:
            line = line[:pos]
        out.append(line)
    return "\t".join(out)


def load_camera(path: str, max_intensity: float = 255.0) -> PinholeCamera:
    with open(path) as f:
        st = _Stream(_strip_comments(f.read()))
    kind = st.word()
    if kind != "PINHOLE":
        raise ValueError(f"{path}: only the PINHOLE camera model is supported, got {kind}")
    st.word()
    st.word()
    n_row, n_col = (int(v) for v in st.word().split(",")[:2])
    K = np.array([st.number() for _ in range(9)]).reshape(3, 3)
    dist = [float(v) for v in st.word().split(",") if v.strip()]
    st.word()  # rotation vector (not used)
    R = np.array([st.number() for _ in range(9)]).reshape(3, 3)
    Rinv = np.array([st.number() for _ in range(9)]).reshape(3, 3)
    T = np.array([st.number() for _ in range(3)])
    Tinv = np.array([st.number() for _ in range(3)])
    d12 = np.zeros(12)
    d12[: len(dist)] = dist
    cam = PinholeCamera(n_row=n_row, n_col=n_col, fx=K[0, 0], fy=K[1, 1], cx=K[0, 2],
                        cy=K[1, 2], dist=d12, n_dist=len(dist), R=R, T=T, Rinv=Rinv,
                        Tinv=Tinv, max_intensity=max_intensity)
    if cam.is_distorted and cam.n_dist not in (4, 5, 8, 12):
        raise ValueError(f"{path}: number of distortion coefficients is wrong: {cam.n_dist}")
    return cam


def save_camera(cam: PinholeCamera, path: str) -> None:
    def row(v):
        return ",".join(repr(float(x)) for x in v)
    lines = [
        "# Camera Model: (PINHOLE/POLYNOMIAL)", "PINHOLE",
        "# Camera Calibration Error: ", "None", "# Pose Calibration Error: ", "None",
        "# Image Size: (n_row,n_col)", f"{cam.n_row},{cam.n_col}",
        "# Camera Matrix: ", row([cam.fx, 0.0, cam.cx]), row([0.0, cam.fy, cam.cy]),
        row([0.0, 0.0, 1.0]),
        "# Distortion Coefficients: ", row(cam.dist[: max(cam.n_dist, 1)]),
        "# Rotation Vector: ", row([0.0, 0.0, 0.0]),
        "# Rotation Matrix: ", *[row(r) for r in cam.R],
        "# Inverse of Rotation Matrix: ", *[row(r) for r in cam.Rinv],
        "# Translation Vector: ", row(cam.T),
        "# Inverse of Translation Vector: ", row(cam.Tinv), "",
    ]
    with open(path, "w") as f:
        f.write("\n".join(lines))
