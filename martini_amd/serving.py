"""Serving form: a persistent insertion service for a stream of mock observations.

Requests (one source each, with its own cube shape, kernel, distance and
inclination) arrive over time.  Each request's raw particle arrays are
resident in HBM when it arrives.  The service keeps up to ``slots``
requests on the GPU at once, each on its own HIP stream, so that their
kernels overlap on the device, and admits a waiting request as soon as a
slot and enough HBM are free (continuous batching, not a queue in front of
one job).  Every request runs:

  1. ``mi_prepare``: FP64 setup on the device (projection, smoothing
     lengths, kernel choice, pruning, records) and the pair count, whose
     four counters are copied to pinned host memory;
  2. once the counters are back: admission of the insertion working set
     (workspace + cube) against the HBM budget;
  3. ``mi_bin`` + ``mi_gather`` into a zeroed cube, and the cube's copy to
     pinned host memory.

Small requests that wait together and share a cube shape are fused: their
records address consecutive channel ranges of one stacked cube, so a single
bin and gather serve them all (``Service`` fuse_max_n / fuse_max_jobs).

Latency is measured from the request's arrival to the host seeing its cube.
A request inserted on its own gets exactly the cube it gets alone (no state
is shared between insertions and the kernels' summation order is fixed); a
fused request's particles are chunked with its neighbours', which changes
only the FP32 rounding (within 1e-5 of the peak).  ``tests/test_serving.py``
checks both.

Requests are independent ``Martini`` insertions: the per-request semantics
of ``Martini(...)`` + ``insert_source_in_cube`` (martini/martini.py:827-861)
with a fresh cube per request, as ``reset`` (:409-425) gives.
"""

from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .datacube import DataCube
from .device_setup import DeviceSource, check_flags, observation
from .sph_kernels import CubicSplineKernel, WendlandC2Kernel
from .spectral_models import GaussianSpectrum

SHAPES = ((64, 32), (128, 64), (256, 128), (512, 256))  # (n_px, n_channels)
KERNELS = {"cubic": CubicSplineKernel, "wc2": WendlandC2Kernel}
BAND_KMS = 600.0  # spectral band of every request's cube
FOV_MARGIN = 1.1  # cube width over the disc's diameter (15 kpc truncation radius)


@dataclass
class RequestSpec:
    """One mock observation: a synthetic disc and how it is observed."""

    rid: int
    n: int
    seed: int
    incl: float  # deg
    distance: float  # Mpc
    n_px: int
    n_channels: int
    kernel: str
    arrival: float = 0.0  # s after the stream starts

    @property
    def px_size(self) -> float:
        theta = math.atan(15.0 / (self.distance * 1000.0)) * 180.0 / math.pi * 3600.0
        return 2.0 * FOV_MARGIN * theta / self.n_px

    @property
    def channel_width(self) -> float:
        return BAND_KMS / self.n_channels

    @property
    def h_median(self) -> float:  # kpc; 0.8 at 1e4 particles, 0.15 at 1e7
        return 0.8 * (self.n / 1e4) ** -0.24

    def datacube(self) -> DataCube:
        return DataCube(n_px_x=self.n_px, n_px_y=self.n_px, n_channels=self.n_channels,
                        px_size=self.px_size, channel_width=self.channel_width,
                        spectral_centre=70.0 * self.distance)


def request_stream(n_requests: int, seed: int = 5000, n_range=(1e4, 1e7), rate_hz: float = 50.0,
                   shapes=SHAPES, kernels=("cubic", "wc2")) -> list[RequestSpec]:
    """Request mix: particle count log-uniform in n_range, inclination U(0, 90)
    deg, distance U(3, 30) Mpc, cube shape and kernel drawn uniformly, Poisson
    arrivals at rate_hz.  Request i's disc seed is seed + i."""
    rng = np.random.default_rng(seed)
    lo, hi = np.log(n_range[0]), np.log(n_range[1])
    t = 0.0
    out = []
    for i in range(n_requests):
        t += rng.exponential(1.0 / rate_hz)
        n_px, nc = shapes[rng.integers(len(shapes))]
        out.append(RequestSpec(rid=i, n=int(np.exp(rng.uniform(lo, hi))), seed=seed + i,
                               incl=float(rng.uniform(0.0, 90.0)),
                               distance=float(rng.uniform(3.0, 30.0)), n_px=int(n_px),
                               n_channels=int(nc), kernel=str(kernels[rng.integers(len(kernels))]),
                               arrival=t))
    return out


def disc_on_device(spec: RequestSpec, device=None) -> DeviceSource:
    """The exponential-disc recipe of martini_amd.synthetic, drawn on the GPU
    (torch generator seeded with spec.seed): R ~ Gamma(2, 3 kpc) cut at
    15 kpc, z ~ N(0, 0.3 kpc), circular speed 200 tanh(R / 2 kpc) km/s plus
    N(0, 10 km/s) per component, masses U(0.5, 1.5) x 1e9 / N Msun, h
    lognormal (median h_median, sigma_ln 0.5), T = 1e4 K; face-on."""
    import torch

    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else device
    g = torch.Generator(device=dev)
    g.manual_seed(int(spec.seed))
    n = spec.n
    f64 = dict(dtype=torch.float64, device=dev)
    R = torch.empty(0, **f64)
    while R.numel() < n:
        m = int((n - R.numel()) * 1.05) + 64
        u = torch.rand((2, m), generator=g, **f64).clamp_min_(1e-300)
        draw = -3.0 * torch.log(u[0] * u[1])  # Gamma(2, 3): sum of two exponentials
        R = torch.cat((R, draw[draw <= 15.0]))
    R = R[:n]
    phi = torch.rand(n, generator=g, **f64) * (2.0 * math.pi)
    z = torch.randn(n, generator=g, **f64) * 0.3
    noise = torch.randn((n, 3), generator=g, **f64) * 10.0
    vc = 200.0 * torch.tanh(R / 2.0)
    xyz = torch.stack((z, R * torch.cos(phi), R * torch.sin(phi)), dim=1)
    vxyz = torch.stack((noise[:, 0], -vc * torch.sin(phi) + noise[:, 1],
                        vc * torch.cos(phi) + noise[:, 2]), dim=1)
    mHI = (0.5 + torch.rand(n, generator=g, **f64)) * (1e9 / n)
    hsm = spec.h_median * torch.exp(torch.randn(n, generator=g, **f64) * 0.5)
    T = torch.full((n,), 1e4, **f64)
    return DeviceSource(xyz, vxyz, mHI, hsm, T, device=dev)


_KERNEL_OBJECTS: dict = {}


def _kernel(name: str):
    """One kernel object per family (its FWHM root is found once)."""
    if name not in _KERNEL_OBJECTS:
        _KERNEL_OBJECTS[name] = KERNELS[name]()
    return _KERNEL_OBJECTS[name]


def observe(spec: RequestSpec):
    """(mi_observation, mi_cube_desc, datacube) of a request."""
    from scipy.spatial.transform import Rotation

    dc = spec.datacube()
    obs, desc = observation(distance=spec.distance, datacube=dc,
                            sph_kernel=_kernel(spec.kernel),
                            spectral_model=GaussianSpectrum(sigma="thermal"),
                            rotation=Rotation.from_euler("y", np.deg2rad(spec.incl)), h=0.7)
    return obs, desc, dc


@dataclass
class Result:
    rid: int
    arrival: float
    admitted: float = 0.0   # prepare launched
    inserted: float = 0.0   # insertion launched
    done: float = 0.0
    kept: int = 0
    pairs: int = 0
    updates: int = 0        # particle-voxel updates, reference mask semantics
    group: int = 1          # requests inserted together with this one (fusion)
    cube: np.ndarray | None = None

    @property
    def latency(self) -> float:
        return self.done - self.arrival


@dataclass
class _Member:
    spec: RequestSpec
    src: DeviceSource
    res: Result
    obs: object = None
    desc: object = None


@dataclass
class _Job:
    """One insertion on the GPU: a single request, or several small requests of
    one cube shape stacked along the channel axis of a shared cube."""

    members: list
    slot: int = -1
    phase: str = "wait"
    reserved: int = 0
    bufs: dict = field(default_factory=dict)
    event: object = None
    desc: object = None  # descriptor of the (possibly stacked) cube


class Service:
    """Continuous-batching insertion service on one GPU.

    slots: insertions on the GPU at once (one HIP stream each, default priority).
    hbm_budget: bytes of HBM for the working sets (records, setup and insertion
    workspaces, cubes; the raw particle arrays are the requests' own).
    method: 'plain' or 'mfma'.
    fuse_max_n, fuse_max_jobs: requests of at most fuse_max_n particles that
    wait together and share a cube shape are inserted together, up to
    fuse_max_jobs of them, their cubes stacked along the channel axis (each
    record's channel window lies inside its own request's channels, so the
    binning and gather kernels run unchanged on the stacked cube).  One bin and
    one gather then serve all of them.  0 disables fusion.
    """

    def __init__(self, slots: int = 4, hbm_budget: int = 64 << 30, method: str = "mfma",
                 keep_cubes: bool = True, fuse_max_n: int = 100_000,
                 fuse_max_jobs: int = 16) -> None:
        import torch

        self.torch = torch
        self.dev = torch.device("cuda", torch.cuda.current_device())
        self.slots = int(slots)
        self.streams = [torch.cuda.Stream(device=self.dev) for _ in range(self.slots)]
        self.budget = int(hbm_budget)
        self.method = method
        self.keep_cubes = keep_cubes
        self.fuse_max_n = int(fuse_max_n)
        self.fuse_max_jobs = max(1, int(fuse_max_jobs))
        self.in_use = 0
        self.pinned_counts = [torch.zeros((self.fuse_max_jobs, 4), dtype=torch.int64).pin_memory()
                              for _ in range(self.slots)]
        self._pinned_cubes: dict[int, list] = {}
        self.admission_waits = 0
        self.fused_launches = 0

    def warm(self, shapes=SHAPES) -> None:
        """Pins one host cube buffer per slot and shape class up front, so that
        pinned allocation stays out of request latency."""
        for n_px, nc in shapes:
            numel = n_px * n_px * nc
            pool = self._pinned_cubes.setdefault(numel, [])
            while len(pool) < self.slots:
                pool.append(self.torch.empty(numel, dtype=self.torch.float32).pin_memory())

    # ---------------------------------------------------------- buffers
    def _pinned_cube(self, numel: int):
        pool = self._pinned_cubes.setdefault(numel, [])
        if pool:
            return pool.pop()
        return self.torch.empty(numel, dtype=self.torch.float32).pin_memory()

    def _release_pinned(self, t) -> None:
        self._pinned_cubes.setdefault(t.numel(), []).append(t)

    def _reserve(self, nbytes: int, force: bool = False) -> bool:
        if force or self.in_use + nbytes <= self.budget:
            self.in_use += nbytes
            return True
        return False

    # ---------------------------------------------------------- grouping
    def _take_group(self, waiting: deque) -> list:
        """The next job's members: the head request, plus (when it is small)
        the small requests waiting behind it with the same cube and tables."""
        head = waiting.popleft()
        if self.fuse_max_n <= 0 or head.spec.n > self.fuse_max_n or self.fuse_max_jobs < 2:
            return [head]
        head.obs, head.desc, _ = observe(head.spec)
        group, rest = [head], deque()
        while waiting:
            m = waiting.popleft()
            s, h = m.spec, head.spec
            if (len(group) < self.fuse_max_jobs and s.n <= self.fuse_max_n
                    and (s.n_px, s.n_channels) == (h.n_px, h.n_channels)):
                m.obs, m.desc, _ = observe(s)
                if (m.desc.spectrum == head.desc.spectrum
                        and tuple(m.desc.gauss_t) == tuple(head.desc.gauss_t)
                        and tuple(m.desc.gauss_inv_norm) == tuple(head.desc.gauss_inv_norm)):
                    group.append(m)
                    continue
            rest.append(m)
        waiting.extend(rest)
        return group

    # ---------------------------------------------------------- stages
    def _launch_prepare(self, job: _Job, now: float) -> bool:
        torch = self.torch
        B = len(job.members)
        need = sum(48 * max(m.src.n, 1) + _lib.prepare_workspace_bytes(m.src.n) + 256
                   for m in job.members)
        if not self._reserve(need, force=self.in_use == 0):
            return False
        job.reserved = need
        s = self.streams[job.slot]
        counts = torch.empty((B, 4), dtype=torch.int64, device=self.dev)
        outs, wss = [], []
        with torch.cuda.stream(s):
            for b, m in enumerate(job.members):
                if m.obs is None:
                    m.obs, m.desc, _ = observe(m.spec)
                m.obs.channel_offset = b * m.spec.n_channels if B > 1 else 0
                m.obs.nc_total = B * m.spec.n_channels if B > 1 else 0
                n = m.src.n
                out = torch.empty(48 * max(n, 1), dtype=torch.uint8, device=self.dev)
                ws = torch.empty(max(_lib.prepare_workspace_bytes(n), 256), dtype=torch.uint8,
                                 device=self.dev)
                _lib.prepare(m.src.xyz, m.src.vxyz, m.src.mHI, m.src.hsm, m.src.T, None, n, m.obs,
                             out, counts[b], ws, s)
                outs.append(out)
                wss.append(ws)
                m.res.admitted = now
            self.pinned_counts[job.slot][:B].copy_(counts, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(s)
        d0 = job.members[0].desc
        job.desc = _lib.make_desc(d0.nx, d0.ny, d0.nc * B, d0.spectrum)
        for k in range(4):
            job.desc.gauss_t[k], job.desc.gauss_inv_norm[k] = d0.gauss_t[k], d0.gauss_inv_norm[k]
        job.bufs = {"outs": outs, "prep_ws": wss, "counts": counts}
        job.event = ev
        job.phase = "prepare"
        return True

    def _launch_insert(self, job: _Job, now: float) -> bool:
        torch = self.torch
        B = len(job.members)
        rows = [[int(v) for v in self.pinned_counts[job.slot][b]] for b in range(B)]
        for _, _, flags, _ in rows:
            check_flags(flags, skip_validation=True)
        kept = sum(r[0] for r in rows)
        pairs = sum(r[1] for r in rows)
        d = job.desc
        ws_bytes = _lib.workspace_bytes(d, kept, pairs)
        cube_n = d.nx * d.ny * d.nc
        need = ws_bytes + 4 * cube_n + (48 * kept if B > 1 else 0)
        prep_ws = sum(w.numel() for w in job.bufs["prep_ws"])
        if not self._reserve(max(0, need - prep_ws), force=self.in_use - job.reserved == 0):
            self.admission_waits += job.phase != "wait_mem"
            job.phase = "wait_mem"
            return False
        job.reserved += max(0, need - prep_ws)
        s = self.streams[job.slot]
        with torch.cuda.stream(s):
            job.bufs.pop("prep_ws")
            if B == 1:
                rec = job.bufs["outs"][0]
            else:  # the members' records, back to back
                rec = torch.empty(48 * max(kept, 1), dtype=torch.uint8, device=self.dev)
                at = 0
                for out, r in zip(job.bufs["outs"], rows):
                    if r[0]:
                        rec[at:at + 48 * r[0]].copy_(out[:48 * r[0]])
                    at += 48 * r[0]
                self.fused_launches += 1
            ws = torch.empty(max(ws_bytes, 256), dtype=torch.uint8, device=self.dev)
            cube = torch.zeros(cube_n, dtype=torch.float32, device=self.dev)
            if kept > 0:
                _lib.bin(rec, kept, pairs, d, ws, s, method=self.method)
                _lib.gather(rec, kept, pairs, d, cube, ws, self.method, s)
            host = self._pinned_cube(cube_n)
            host.copy_(cube, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(s)
        job.bufs.update(rec=rec, ws=ws, cube=cube, host=host)
        job.event = ev
        job.phase = "insert"
        for m, r in zip(job.members, rows):
            m.res.inserted = now
            m.res.kept, m.res.pairs = r[0], r[1]
            m.res.updates = r[3] * m.spec.n_channels
            m.res.group = B
        return True

    def _finish(self, job: _Job, now: float) -> None:
        host = job.bufs["host"]
        d = job.desc
        B = len(job.members)
        if self.keep_cubes:
            full = host.numpy().reshape(d.nx, d.ny, B, d.nc // B)
            for b, m in enumerate(job.members):
                m.res.cube = full[:, :, b, :].copy()
        self._release_pinned(host)
        job.bufs.clear()
        self.in_use -= job.reserved
        for m in job.members:
            m.res.done = now
        job.phase = "done"

    # ---------------------------------------------------------- loop
    def run(self, specs, sources) -> list[Result]:
        """Serve ``specs`` (sorted by arrival) whose raw arrays are ``sources``
        (DeviceSource per request, already in HBM).  Returns per-request results;
        times are seconds after the start of the stream."""
        torch = self.torch
        torch.cuda.synchronize()
        pending = deque(sorted(zip(specs, sources), key=lambda p: p[0].arrival))
        waiting: deque[_Member] = deque()
        active: dict[int, _Job] = {}
        results: list[Result] = []
        t0 = time.perf_counter()
        while pending or waiting or active:
            now = time.perf_counter() - t0
            while pending and pending[0][0].arrival <= now:
                spec, src = pending.popleft()
                waiting.append(_Member(spec, src, Result(spec.rid, spec.arrival)))
            progressed = False
            for slot, job in list(active.items()):
                if job.phase == "prepare" and job.event.query():
                    progressed |= self._launch_insert(job, now)
                elif job.phase == "wait_mem":
                    progressed |= self._launch_insert(job, now)
                elif job.phase == "insert" and job.event.query():
                    self._finish(job, time.perf_counter() - t0)
                    results.extend(m.res for m in job.members)
                    del active[slot]
                    progressed = True
            free = [k for k in range(self.slots) if k not in active]
            while waiting and free:
                group = self._take_group(waiting)
                job = _Job(group, slot=free[0])
                if not self._launch_prepare(job, now):
                    for m in reversed(group):
                        waiting.appendleft(m)
                    break
                active[free.pop(0)] = job
                progressed = True
            if not progressed:
                if pending and not waiting and not active:
                    time.sleep(max(0.0, min(pending[0][0].arrival - now, 1e-3)))
                else:
                    time.sleep(2e-5)
        torch.cuda.synchronize()
        return sorted(results, key=lambda r: r.rid)


def summarize(results: list[Result], wall: float | None = None) -> dict:
    lat = np.array([r.latency for r in results]) * 1e3
    queue = np.array([r.admitted - r.arrival for r in results]) * 1e3
    upd = float(sum(r.updates for r in results))
    if wall is None:
        wall = max(r.done for r in results) - min(r.arrival for r in results)
    return {
        "requests": len(results),
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "mean_ms": float(lat.mean()),
        "queue_p50_ms": float(np.percentile(queue, 50)),
        "wall_s": float(wall),
        "updates": upd,
        "updates_per_s": upd / wall,
        "particles": int(sum(r.kept for r in results)),
        "fused_requests": int(sum(r.group > 1 for r in results)),
    }
