"""RMSE-versus-time harness: CUDAPTWindow::CalculateRMSE and its capture /
data-collection loop (src/cudapt.cpp:268-337,951-965).

The reference renders one frame per window iteration and, while capturing,
adds the frame's wall time and counts frames; it computes the RMSE against a
captured converged image either once (after a time or frame budget, or when
the RMSE falls to a limit) or every `interval` seconds, writing the collected
rows as text lines "sec, frames, rmse".

Here a "frame" is one render call of `frames_per_step` frames (1 = the
reference), the time is the synchronised wall time of the render calls only
(the RMSE readback and sum are outside it, as the reference intends at
cudapt.cpp:300-301,329), and the RMSE is the reference's single-precision
sequential sum over W*H*3 floats.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np


def rmse(converged, image) -> float:
    """sqrt(sum((conv - img)^2) / (W*H*3)) with the sum accumulated in float in
    index order (cudapt.cpp:951-965)."""
    c = np.ascontiguousarray(converged, np.float32).ravel()
    i = np.ascontiguousarray(image, np.float32).ravel()
    if c.size != i.size:
        raise ValueError(f"image sizes differ: {c.size} vs {i.size}")
    if c.size == 0:
        return 0.0
    d = c - i
    total = np.cumsum(d * d, dtype=np.float32)[-1]
    return float(np.sqrt(np.float32(total) / np.float32(c.size), dtype=np.float32))


@dataclass
class Sample:
    sec: float
    frames: int
    rmse: float


@dataclass
class RmseHarness:
    """Drives `renderer.render(scene, camera, frames=k)` and records RMSE rows.

    `renderer` is a cudapt_amd.Renderer (or anything with the same render()
    signature returning the running mean as a H x W x 3 float array)."""
    renderer: object
    scene: object
    camera: dict
    converged: np.ndarray
    frames_per_step: int = 1
    rows: list = field(default_factory=list)
    elapsed: float = 0.0
    frames: int = 0
    _image: np.ndarray | None = None

    def step(self) -> np.ndarray:
        t0 = time.perf_counter()
        img = self.renderer.render(self.scene, self.camera, frames=self.frames_per_step)
        self.elapsed += time.perf_counter() - t0
        self.frames += self.frames_per_step
        self._image = img
        return img

    def current_rmse(self) -> float:
        if self._image is None:
            raise RuntimeError("nothing rendered yet")
        return rmse(self.converged, self._image)

    def capture(self, seconds: float = 0.0, frames: int = 0, limit: float = 0.0, max_frames: int = 1 << 20) -> Sample:
        """One RMSE measurement (cudapt.cpp:286-323): after `seconds` of rendering,
        after `frames` frames, or at the first frame whose RMSE is <= `limit`."""
        if not (seconds > 0 or frames > 0 or limit > 0):
            raise ValueError("give a time, frame or RMSE budget")
        while True:
            self.step()
            done = (seconds > 0 and self.elapsed >= seconds) or (frames > 0 and self.frames >= frames)
            if limit > 0:
                e = self.current_rmse()
                if 0 < e <= limit or self.frames >= max_frames:
                    return Sample(self.elapsed, self.frames, e)
            elif done or self.frames >= max_frames:
                return Sample(self.elapsed, self.frames, self.current_rmse())

    def collect(self, interval: float, count: int, max_frames: int = 1 << 20) -> list:
        """Data collection (cudapt.cpp:279-307): a row every `interval` seconds of
        rendering time, `count` rows."""
        if interval <= 0 or count < 1:
            raise ValueError("interval must be > 0 and count >= 1")
        n = 0
        while n < count and self.frames < max_frames:
            self.step()
            if self.elapsed >= (n + 1) * interval:
                n += 1
                self.rows.append(Sample(self.elapsed, self.frames, self.current_rmse()))
        return self.rows

    def csv(self) -> str:
        """The collected rows as the reference writes them: "sec, frames, rmse"."""
        return "".join(f"{r.sec}, {r.frames}, {r.rmse}\n" for r in self.rows)

    def write_csv(self, path: str):
        with open(path, "w") as f:
            f.write(self.csv())


def read_csv(path: str) -> list:
    rows = []
    with open(path) as f:
        for line in f:
            if line.strip():
                s, fr, e = (x.strip() for x in line.split(","))
                rows.append(Sample(float(s), int(float(fr)), float(e)))
    return rows
