"""Continuous-batching FBP / FDK server on one device.

Jobs (raw FP32 projections + a FanGeometry or ConeGeometry) arrive over time. The server

* admits a job once its HBM footprint fits the remaining budget and an active slot is free
  (first come, first admitted);
* prepares it on one of two prep streams (jobs with little filtering work on their own, so
  they do not wait behind a large job's filtering): cosine weights, ramp filter (row pairs through complex
  torch.fft transforms, rocFFT) in FP32, scale, and staging of the filtered projections into
  the zero-bordered layout of the backprojector in the storage precision (FP16 by default);
  the volume is allocated in the storage precision and stays resident;
* backprojects in rounds: each round is one persistent launch of the native multi-job
  backprojector (cmi_serve_backproject) over tiles of every prepared job, at most
  `round_cost` pixel-angle units, shortest remaining work first with a guaranteed share for
  the oldest job, so small jobs are not stuck behind large ones and large ones keep moving;
  up to `max_rounds_in_flight` rounds are queued, alternating between compute streams so
  that one launch's workgroups fill the compute units another launch frees as it drains,
  while new arrivals are prepared on the prep stream, so filtering overlaps backprojection;
* completes a job when the round holding its last tiles has finished, releasing its staged
  projections and keeping its volume as the result.

Each voxel is computed by one thread with the plain backprojector's arithmetic, so a job's
result is the same alone or in any mix, and the same as `curadon_amd.fbp` with the same
storage precision.
"""
from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np
import torch

from . import _lib
from .geometry import ConeGeometry, FanGeometry
from .reconstruction import (cosine_weights_cone, cosine_weights_fan, fft_length,
                             filter_rows_weighted)

# Relative cost of one voxel-angle of the cone-beam backprojector to one pixel-angle of the
# fan-beam one (2 x 2 taps from LDS-staged footprints against 2 taps from the sinogram; measured
# 2.1e12 voxel-angles/s at 384^3 against 2.35e12 pixel-angles/s at 2048^2).
CONE_COST = 1.15
PIXELS_PER_TILE_2D = 64 * 4
# voxels per tile of the cone-beam kernel: y runs (16 x 16 x 16) / z columns (16 x 16 x 8)
VOXELS_PER_TILE_3D = {True: 16 * 16 * 16, False: 16 * 16 * 8}


# Projection elements (rows x cells) up to which a job's filtering counts as small and is
# prepared on the small-job prep stream.
PREP_SMALL_ELEMENTS = 1 << 22


def _prep_elements(g) -> int:
    if isinstance(g, FanGeometry):
        return int(g.nangles) * int(g.det_shape)
    return int(g.nangles) * int(np.prod([int(v) for v in g.det_shape]))


@dataclass
class FBPJob:
    """One reconstruction request. `projections` are raw FP32 projections, (A, D) for a fan
    beam or (A, V, U) for a cone beam, on the server's device or on the host."""
    geom: object
    projections: torch.Tensor
    window: str = "ram-lak"
    frequency_scaling: float = 1.0
    arrival: float = 0.0
    tag: str = ""
    # filled in by the server
    id: int = -1
    volume: Optional[torch.Tensor] = None
    t_admit: float = math.nan
    t_done: float = math.nan

    @property
    def latency(self) -> float:
        return self.t_done - self.arrival


@dataclass
class _Active:
    job: FBPJob
    fan: bool
    ntiles: int
    tile_cost: float
    footprint: int
    staged: torch.Tensor
    table: torch.Tensor
    ycols: bool = False
    keep: list = field(default_factory=list)
    prep_done: Optional[torch.cuda.Event] = None
    next_tile: int = 0
    tiles_done: int = 0

    @property
    def remaining_cost(self) -> float:
        return (self.ntiles - self.next_tile) * self.tile_cost


def footprint(job: FBPJob, storage: torch.dtype):
    """(resident, transient) HBM bytes of a job. Resident, held while the job is active:
    staged projections, table and volume. Transient, held while it is prepared: the FP32
    filtering buffers (packed row pairs, their spectrum, the filtered pairs and the filtered
    rows; at most two of the first three at once, bounded here by the three), and the raw
    projections when they come from the host."""
    s = torch.tensor([], dtype=storage).element_size()
    g = job.geom
    if isinstance(g, FanGeometry):
        A, D = g.nangles, g.det_shape
        ny, nx = (int(v) for v in g.vol_shape)
        L, rows, U = fft_length(D), A, D
        vox, staged, table = nx * ny, A * (D + 4) * s, 16 * A
    else:
        A = g.nangles
        V, U = (int(v) for v in g.det_shape)
        L, rows = fft_length(U), A * V
        vox, staged, table = int(np.prod(g.vol_shape)), A * (V + 4) * (U + 4) * s, 48 * A
    transient = rows * U * 4 + 2 * rows * (L // 2 + 1) * 8 + rows * L * 4
    if not job.projections.is_cuda:
        transient += rows * U * 4
    return int(staged + vox * s + table), int(transient)


class FBPServer:
    def __init__(self, device: int = 0, storage: torch.dtype = torch.float16,
                 hbm_budget: int = 32 << 30, max_active: int = 64,
                 round_cost: float = 4e9, oldest_share: float = 0.5,
                 max_rounds_in_flight: int = 3, compute_streams: int = 2):
        if storage not in (torch.float16, torch.bfloat16, torch.float32):
            raise TypeError("storage must be float16, bfloat16 or float32")
        self.device = torch.device("cuda", device)
        self.storage = storage
        self.hbm_budget = int(hbm_budget)
        self.max_active = int(max_active)
        self.round_cost = float(round_cost)
        self.oldest_share = float(oldest_share)
        self.max_rounds_in_flight = int(max_rounds_in_flight)
        self._native = _lib.Server(device, max_jobs=max_active)
        # two prep streams (default priority): jobs with little filtering work are prepared on
        # their own stream, so they do not queue behind a large job's filtering
        self._prep_small = torch.cuda.Stream(self.device)
        self._prep_large = torch.cuda.Stream(self.device)
        # launches alternate between streams: while one launch drains its last tiles, the
        # next one's workgroups take the freed compute units
        self._compute = [torch.cuda.Stream(self.device) for _ in range(max(1, compute_streams))]
        self._cache = {}
        self.rounds = 0
        self.launch_log: List[tuple] = []
        self.max_active_seen = 0
        self.max_used_seen = 0

    # ------------------------------------------------------------------ prep ----------
    def _filter_parts(self, g, window, fs):
        # keyed by the geometry's attributes, so a geometry changed after an earlier job gets
        # new weights and tables
        key = (type(g).__name__, g.fingerprint(), window, fs)
        hit = self._cache.get(key)
        if hit is not None:
            return hit
        dev = self.device
        if isinstance(g, FanGeometry):
            cw = cosine_weights_fan(g, dev)
            scale = ((math.pi / g.nangles) * torch.as_tensor(
                g.DSO / g.DSD, dtype=torch.float32, device=dev)).contiguous()
            table = torch.from_numpy(_lib.fan2d_bp_table(g.angles, g.DSO, g.DSD)).to(dev)
            parts = (cw, scale, 1.0, table)
        else:
            cw = cosine_weights_cone(g, dev)
            scale = (math.pi / g.nangles) * g.DSD / (float(g.det_spacing[1]) * g.DSO)
            # host-side matrices: no device plan, no synchronisation on the serving loop
            m, yinv = g.bp_matrices("fdk", with_layout=True)
            table = torch.from_numpy(m).to(dev)
            parts = (cw, None, scale, (table, yinv))
        if len(self._cache) > 256:
            self._cache.clear()
        self._cache[key] = parts
        return parts

    def _prepare(self, job: FBPJob) -> _Active:
        g = job.geom
        fan = isinstance(g, FanGeometry)
        if not fan and not isinstance(g, ConeGeometry):
            raise ValueError("Unknown geometry type")
        dev = self.device
        prep = self._prep_large if _prep_elements(g) > PREP_SMALL_ELEMENTS else self._prep_small
        with torch.cuda.stream(prep):
            cw, row_scale, scale, table = self._filter_parts(g, job.window,
                                                             job.frequency_scaling)
            ycols = False
            if not fan:
                table, ycols = table
            p = job.projections.to(device=dev, dtype=torch.float32, non_blocking=True)
            if fan:
                A, D = g.nangles, g.det_shape
                if tuple(p.shape) != (A, D):
                    raise ValueError(f"projections must be {(A, D)}, got {tuple(p.shape)}")
                ny, nx = (int(v) for v in g.vol_shape)
                y = filter_rows_weighted(p, cw, D, job.window, job.frequency_scaling)
                staged = torch.empty((A, D + 4), dtype=self.storage, device=dev)
                keep = [_lib.stage_projections(y, staged, row_scale=row_scale)]
                volume = torch.zeros((ny, nx), dtype=self.storage, device=dev)
                ntiles = _lib.bp_job_tiles(_lib.BPJob(kind=_lib.JOB_FAN2D, nx=nx, ny=ny))
                tile_cost = float(PIXELS_PER_TILE_2D * A)
            else:
                A = g.nangles
                V, U = (int(v) for v in g.det_shape)
                if tuple(p.shape) != (A, V, U):
                    raise ValueError(f"projections must be {(A, V, U)}, got {tuple(p.shape)}")
                nz, ny, nx = (int(v) for v in g.vol_shape)
                y = filter_rows_weighted(p, cw, U, job.window, job.frequency_scaling)
                staged = torch.empty((A, V + 4, U + 4), dtype=self.storage, device=dev)
                keep = [_lib.stage_projections(y, staged, scale=scale)]
                volume = torch.zeros((nz, ny, nx), dtype=self.storage, device=dev)
                ntiles = _lib.bp_job_tiles(_lib.BPJob(kind=_lib.JOB_CONE3D, columns=int(ycols),
                                                      nx=nx, ny=ny, nz=nz))
                tile_cost = float(VOXELS_PER_TILE_3D[bool(ycols)] * A * CONE_COST)
            ev = torch.cuda.Event()
            ev.record(prep)
        job.volume = volume
        # p and y are released here: the allocator reuses their blocks on the same prep stream
        # only, after the staging kernel in stream order.
        return _Active(job=job, fan=fan, ntiles=ntiles, tile_cost=tile_cost,
                       footprint=footprint(job, self.storage)[0], staged=staged, table=table,
                       ycols=ycols, keep=keep, prep_done=ev)

    def _descriptor(self, a: _Active, t0: int, t1: int) -> _lib.BPJob:
        g, j = a.job.geom, _lib.BPJob()
        j.sino, j.table, j.volume = a.staged.data_ptr(), a.table.data_ptr(), a.job.volume.data_ptr()
        j.tile_begin, j.tile_end = t0, t1
        if a.fan:
            j.kind, j.mode = _lib.JOB_FAN2D, _lib.BP2D_FBP
            ny, nx = (int(v) for v in g.vol_shape)
            j.nx, j.ny, j.nz = nx, ny, 1
            j.nu, j.nv, j.nangles = g.det_shape, 1, g.nangles
            j.sx, j.sy = float(g.vol_spacing[1]), float(g.vol_spacing[0])
            j.ox, j.oy = float(g.vol_offset[1]), float(g.vol_offset[0])
            j.inv_du = 1.0 / float(g.det_spacing)
        else:
            j.kind, j.mode, j.columns = _lib.JOB_CONE3D, _lib.BP3D_FDK, int(a.ycols)
            nz, ny, nx = (int(v) for v in g.vol_shape)
            V, U = (int(v) for v in g.det_shape)
            j.nx, j.ny, j.nz = nx, ny, nz
            j.nu, j.nv, j.nangles = U, V, g.nangles
        return j

    # ------------------------------------------------------------------ loop ----------
    def run(self, jobs: List[FBPJob], clock=time.perf_counter, poll_sleep: float = 0.0):
        """Serve `jobs` (arrival times in seconds from the start of the call) to completion.
        Returns a dict of throughput / latency statistics; each job gets `volume`, `t_admit`
        and `t_done` on the same clock as `arrival`."""
        pending = deque(sorted(jobs, key=lambda j: j.arrival))
        for i, j in enumerate(jobs):
            if j.id < 0:
                j.id = i
            if getattr(j.geom, "vecs", None) is not None or (
                    isinstance(j.geom, FanGeometry) and not j.geom.is_centred_circular()):
                raise NotImplementedError(
                    f"job {j.id}: geometries from ASTRA-like views (from_vec) or with an exact "
                    f"detector offset, rotation or COR are not served")
            need = sum(footprint(j, self.storage))
            if need > self.hbm_budget:
                raise MemoryError(f"job {j.id} needs {need} B, more than the server's budget "
                                  f"of {self.hbm_budget} B")
        active: List[_Active] = []
        rounds = []  # (event, [(active, tiles)])
        used = 0
        done = 0
        t0 = clock()
        while done < len(jobs):
            now = clock() - t0
            progressed = False
            # admission, first come first admitted
            while (pending and pending[0].arrival <= now and len(active) < self.max_active and
                   used + sum(footprint(pending[0], self.storage)) <= self.hbm_budget):
                j = pending.popleft()
                j.t_admit = now
                a = self._prepare(j)
                used += a.footprint
                active.append(a)
                self.max_active_seen = max(self.max_active_seen, len(active))
                self.max_used_seen = max(self.max_used_seen, used)
                progressed = True
            # completed rounds (launches on different streams may finish out of order)
            for r in [r for r in rounds if r[0].query()]:
                rounds.remove(r)
                parts = r[1]
                now = clock() - t0
                for a, n in parts:
                    a.tiles_done += n
                    if a.tiles_done == a.ntiles:
                        a.job.t_done = now
                        used -= a.footprint
                        active.remove(a)
                        a.staged = a.table = None
                        a.keep.clear()
                        done += 1
                progressed = True
            # next round
            if len(rounds) < self.max_rounds_in_flight:
                parts = self._plan_round(active)
                if parts:
                    stream = self._compute[self.rounds % len(self._compute)]
                    self._launch(parts, stream)
                    ev = torch.cuda.Event()
                    ev.record(stream)
                    rounds.append((ev, [(a, t1 - t0_) for a, t0_, t1 in parts]))
                    progressed = True
            if not progressed:
                if rounds:
                    if poll_sleep:
                        time.sleep(poll_sleep)
                elif pending and not active:
                    wait = pending[0].arrival - (clock() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 1e-3))
                elif active:
                    # prepared jobs not ready yet: wait for the oldest prep
                    for a in active:
                        if a.next_tile < a.ntiles and not a.prep_done.query():
                            a.prep_done.synchronize()
                            break
        torch.cuda.synchronize(self.device)
        return self.stats(jobs)

    def _plan_round(self, active: List[_Active]):
        ready = [a for a in active if a.next_tile < a.ntiles and a.prep_done.query()]
        if not ready:
            return []
        budget = self.round_cost
        parts = {}
        if self.oldest_share > 0:
            # guaranteed share of the oldest job
            oldest = min(ready, key=lambda a: (a.job.arrival, a.job.id))
            n = min(oldest.ntiles - oldest.next_tile,
                    max(1, int(self.oldest_share * budget / oldest.tile_cost)))
            parts[id(oldest)] = (oldest, oldest.next_tile, oldest.next_tile + n)
            budget -= n * oldest.tile_cost
        # shortest remaining work first
        for a in sorted(ready, key=lambda a: (a.remaining_cost, a.job.arrival)):
            if budget <= 0:
                break
            start = parts[id(a)][2] if id(a) in parts else a.next_tile
            left = a.ntiles - start
            if left <= 0:
                continue
            n = min(left, max(1, int(budget / a.tile_cost)))
            begin = parts[id(a)][1] if id(a) in parts else a.next_tile
            parts[id(a)] = (a, begin, start + n)
            budget -= n * a.tile_cost
        out = list(parts.values())
        for a, _, t1 in out:
            a.next_tile = t1
        # costliest tiles first, cheap ones fill the tail of the launch
        out.sort(key=lambda p: -p[0].tile_cost)
        return out

    def _launch(self, parts, stream):
        for a, _, _ in parts:
            stream.wait_event(a.prep_done)
        descs = [self._descriptor(a, t0, t1) for a, t0, t1 in parts]
        self._native.backproject(descs, self.storage, self.storage, stream=stream.cuda_stream)
        self.rounds += 1
        self.launch_log.append(tuple((a.job.id, t0, t1) for a, t0, t1 in parts))

    @staticmethod
    def stats(jobs: List[FBPJob]) -> dict:
        lat = np.array([j.latency for j in jobs]) * 1e3
        t_first = min(j.arrival for j in jobs)
        t_last = max(j.t_done for j in jobs)
        return {
            "jobs": len(jobs),
            "makespan_s": t_last - t_first,
            "jobs_per_s": len(jobs) / (t_last - t_first),
            "latency_ms_p50": float(np.percentile(lat, 50)),
            "latency_ms_p99": float(np.percentile(lat, 99)),
            "latency_ms_mean": float(lat.mean()),
            "latency_ms_max": float(lat.max()),
        }
