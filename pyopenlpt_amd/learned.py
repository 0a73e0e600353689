"""Learned shake proposal: the BF16 MLP (olpt_mlp_* in include/olpt_c.h).

No counterpart in the reference.  ShakeNet wraps one olpt_mlp handle: its
parameters (FP32 master copy, He-initialised here), inference and Adam
training steps run as HIP kernels on MFMA; torch only provides device
buffers and the stream.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check_mlp

N_IN, K0, H, N_OUT, N4 = 484, 512, 256, 4, 32


def param_layout() -> dict:
    """Name -> (offset, shape) in the flat FP32 parameter array."""
    out, off = {}, 0
    for name, shape in (("W1", (K0, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)),
                        ("W3", (H, H)), ("b3", (H,)), ("W4", (H, N4)), ("b4", (N4,))):
        out[name] = (off, shape)
        off += int(np.prod(shape))
    return out


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal hidden weights, a small output layer, zero biases; padding
    (rows 484.. of W1, columns 4.. of W4) zero."""
    rng = np.random.default_rng(seed)
    lay = param_layout()
    p = np.zeros(sum(int(np.prod(s)) for _, s in lay.values()), np.float32)

    def put(name, arr):
        o, s = lay[name]
        p[o:o + arr.size] = arr.reshape(-1)

    w1 = np.zeros((K0, H), np.float32)
    w1[:N_IN] = rng.normal(0, np.sqrt(2.0 / N_IN), (N_IN, H))
    put("W1", w1)
    put("W2", rng.normal(0, np.sqrt(2.0 / H), (H, H)).astype(np.float32))
    put("W3", rng.normal(0, np.sqrt(2.0 / H), (H, H)).astype(np.float32))
    w4 = np.zeros((H, N4), np.float32)
    w4[:, :N_OUT] = rng.normal(0, 0.1 * np.sqrt(1.0 / H), (H, N_OUT))
    put("W4", w4)
    return p


class ShakeNet:
    """One olpt_mlp handle on torch's current stream of `device`."""

    def __init__(self, max_batch: int = 8192, device: int = 0, seed: int = 0,
                 stream: torch.cuda.Stream | None = None):
        self.L = _lib.lib()
        self.n_params = int(self.L.olpt_mlp_num_params())
        self.max_batch = max_batch
        self.stream = stream if stream is not None else torch.cuda.current_stream(device)
        h = ctypes.c_void_p()
        check_mlp(self.L.olpt_mlp_create(device, ctypes.c_void_p(self.stream.cuda_stream),
                                         max_batch, ctypes.byref(h)))
        self.h = h
        self.set_params(init_params(seed))

    def close(self):
        if getattr(self, "h", None):
            self.L.olpt_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _fp(self, a: np.ndarray):
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        check_mlp(self.L.olpt_mlp_set_params(self.h, self._fp(p), p.size))

    def params(self) -> np.ndarray:
        p = np.zeros(self.n_params, np.float32)
        check_mlp(self.L.olpt_mlp_get_params(self.h, self._fp(p), p.size))
        return p

    def grads(self) -> np.ndarray:
        g = np.zeros(self.n_params, np.float32)
        check_mlp(self.L.olpt_mlp_get_grads(self.h, self._fp(g), g.size))
        return g

    def infer(self, x: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """x [n][512] bfloat16 (device) -> [n][4] float32."""
        assert x.dtype == torch.bfloat16 and x.shape[1] == K0 and x.is_contiguous()
        n = x.shape[0]
        if out is None:
            out = torch.empty((n, N_OUT), dtype=torch.float32, device=x.device)
        check_mlp(self.L.olpt_mlp_infer(self.h, ctypes.c_void_p(x.data_ptr()), n,
                                        ctypes.c_void_p(out.data_ptr())))
        return out

    def train_step(self, x: torch.Tensor, t: torch.Tensor, lr: float,
                   ghost_weight: float = 1.0) -> float:
        """One Adam step; t [n][4] = (dx, dy, dz in voxels, ghost 0/1)."""
        assert x.dtype == torch.bfloat16 and x.shape[1] == K0 and x.is_contiguous()
        assert t.dtype == torch.float32 and t.shape == (x.shape[0], 4) and t.is_contiguous()
        loss = ctypes.c_float()
        check_mlp(self.L.olpt_mlp_train_step(self.h, ctypes.c_void_p(x.data_ptr()),
                                             ctypes.c_void_p(t.data_ptr()), x.shape[0],
                                             float(lr), float(ghost_weight),
                                             ctypes.byref(loss)))
        return float(loss.value)

    def save(self, path: str):
        np.savez_compressed(path, params=self.params())

    def load(self, path: str):
        with np.load(path, allow_pickle=False) as z:
            self.set_params(z["params"])


# ---------------------------------------------------------------------------
# training data: frames drawn by the projection kernel
# ---------------------------------------------------------------------------
def training_frame(ctx, params, rng: np.random.Generator, vox: float,
                   ppp_range=(0.01, 0.08), jitter_vox: float = 0.5, noise: float = 10.0):
    """One synthetic frame rendered on the GPU (olpt_render, 5x5 boxes,
    amplitudes U(800, 1200), Gaussian noise, clip, truncate), matched by the
    library's peaks + stereo; every matched object becomes a sample.

    Objects within 1 voxel of a true particle are real: their input position
    is moved by U(-jitter, jitter) voxels per axis and the target is the
    offset back to the truth (voxels).  The others are ghosts (target 0,
    label 1).  Returns (X bf16 [n][512], T float32 [n][4], n_real, n_truth).
    """
    from scipy.spatial import cKDTree
    from . import synth as S
    from ._lib import OLPT_F32
    H, W = ctx.H, ctx.W
    n_truth = int(rng.uniform(*ppp_range) * H * W)
    seed = int(rng.integers(1 << 30))
    truth = (S.clustered_particles(n_truth, seed) if rng.random() < 0.5
             else S.uniform_particles(n_truth, seed))
    amp = rng.uniform(800.0, 1200.0, n_truth)
    dev = torch.device("cuda", ctx.device)
    d_xyz = torch.from_numpy(np.ascontiguousarray(truth, np.float64)).to(dev)
    d_amp = torch.from_numpy(amp).to(dev)
    img = ctx.render(d_xyz, d_amp, r_px=2.5).float()
    g = torch.Generator(device=dev).manual_seed(seed)
    img = torch.floor(torch.clamp(img + noise * torch.randn(img.shape, device=dev, generator=g),
                                  0.0, 65535.0)).contiguous()
    pts = [ctx.find_peaks(img[k].contiguous(), OLPT_F32, params) for k in range(ctx.n_cam)]
    obj = ctx.stereo_match(pts, params, candidates=False)["obj_xyz"]
    if len(obj) == 0:
        return None
    d, idx = cKDTree(truth).query(obj)
    real = d < vox
    jit = rng.uniform(-jitter_vox, jitter_vox, obj.shape) * vox
    x_in = obj + np.where(real[:, None], jit, 0.0)
    T = np.zeros((len(obj), 4), np.float32)
    T[real, :3] = (truth[idx[real]] - x_in[real]) / vox
    T[~real, 3] = 1.0
    X = ctx.shake_patches(img, OLPT_F32, params,
                          torch.from_numpy(np.ascontiguousarray(x_in)).to(dev))
    return X, torch.from_numpy(T).to(dev), int(real.sum()), n_truth


def make_training_set(ctx, params, vox: float, n_particles: int = 2_000_000, seed: int = 9001,
                      **kw):
    """Frames until n_particles true particles have been drawn.  A frame
    whose pruneMatch the library refuses (percentile or e_check values
    within the eq_eps tolerance that are not transitive, status 103 - about
    one dense frame in a hundred) is skipped: the training data needs
    matched objects, not that frame."""
    from ._lib import FatalError
    rng = np.random.default_rng(seed)
    xs, ts, n_seen, n_real, skipped = [], [], 0, 0, 0
    while n_seen < n_particles:
        try:
            r = training_frame(ctx, params, rng, vox, **kw)
        except FatalError as e:
            if e.status != 103:
                raise
            skipped += 1
            continue
        if r is None:
            continue
        xs.append(r[0])
        ts.append(r[1])
        n_real += r[2]
        n_seen += r[3]
    return torch.cat(xs), torch.cat(ts), dict(samples=sum(len(x) for x in xs), real=n_real,
                                              particles=n_seen, frames=len(xs),
                                              skipped_frames=skipped)


def train(net: ShakeNet, X: torch.Tensor, T: torch.Tensor, steps: int, batch: int = 8192,
          lr: float = 1e-3, ghost_weight: float = 0.5, seed: int = 0) -> list[float]:
    """Adam on random batches; the learning rate decays linearly to 10%."""
    g = torch.Generator(device=X.device).manual_seed(seed)
    n = X.shape[0]
    losses = []
    for s in range(steps):
        idx = torch.randint(0, n, (min(batch, n),), device=X.device, generator=g)
        xb = X[idx].contiguous()
        tb = T[idx].contiguous()
        rate = lr * (1.0 - 0.9 * s / max(1, steps - 1))
        losses.append(net.train_step(xb, tb, rate, ghost_weight))
    return losses


# ---------------------------------------------------------------------------
# MLP throughput on MFMA (python -m pyopenlpt_amd.learned)
# ---------------------------------------------------------------------------
MACS_PER_ROW = K0 * H + 2 * H * H + H * N4   # padded GEMM shapes the kernels run


def _bench(argv=None) -> int:
    import argparse
    import json
    import time
    ap = argparse.ArgumentParser(description="time the proposal MLP's training steps and "
                                             "inference (BF16 MFMA, FP32 accumulate)")
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--infer-rows", type=int, default=40960)
    args = ap.parse_args(argv)
    dev = torch.device("cuda", 0)
    g = torch.Generator(device=dev).manual_seed(0)
    net = ShakeNet(max_batch=args.batch, seed=0)
    x = torch.zeros((args.batch, K0), device=dev)
    x[:, :N_IN] = torch.rand((args.batch, N_IN), device=dev, generator=g)
    x = x.to(torch.bfloat16).contiguous()
    t = torch.zeros((args.batch, 4), device=dev)
    t[:, :3] = 0.3 * torch.randn((args.batch, 3), device=dev, generator=g)
    t[:, 3] = (torch.rand(args.batch, device=dev, generator=g) < 0.3).float()
    t = t.contiguous()
    for _ in range(10):
        net.train_step(x, t, 1e-4)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        net.train_step(x, t, 1e-4)
    torch.cuda.synchronize()
    train_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    # forward 2 B MACs, backward 4 B MACs less the unused input-gradient GEMM
    train_flop = 6.0 * args.batch * MACS_PER_ROW - 2.0 * args.batch * K0 * H
    xi = torch.zeros((args.infer_rows, K0), device=dev)
    xi[:, :N_IN] = torch.rand((args.infer_rows, N_IN), device=dev, generator=g)
    xi = xi.to(torch.bfloat16).contiguous()
    y = net.infer(xi)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(20):
        net.infer(xi, y)
    e1.record()
    torch.cuda.synchronize()
    infer_ms = e0.elapsed_time(e1) / 20
    infer_flop = 2.0 * args.infer_rows * MACS_PER_ROW
    peak = 2500.0  # dense BF16 TFLOP/s
    out = {"train_ms_per_step": round(train_ms, 4),
           "train_tflops": round(train_flop / train_ms / 1e9, 2),
           "infer_ms": round(infer_ms, 4), "infer_rows": args.infer_rows,
           "infer_tflops": round(infer_flop / infer_ms / 1e9, 2),
           "peak_tflops_bf16_dense": peak, "batch": args.batch}
    out["train_frac"] = round(out["train_tflops"] / peak, 4)
    out["infer_frac"] = round(out["infer_tflops"] / peak, 4)
    print(json.dumps(out), flush=True)
    net.close()
    return 0


if __name__ == "__main__":
    import sys
    sys.exit(_bench())
