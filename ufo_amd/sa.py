"""StableAperture: turns survived by each particle (ufo/sa.py:18-145)."""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import codegen
from .sim import Simulation, host_buffer


class StableAperture(Simulation):
    """Track each particle until it is lost or ``turns`` turns are done.

    Same arguments and attributes as the reference (ufo/sa.py:37-145):
    ``parameters`` is [particles, max(1, len(parameters))]; after ``run()``
    ``lost[p]`` is the number of turns completed before the end-of-turn test
    ``|x| > 1 or |y| > 1 or isnan(x)`` fired, or ``turns`` if the particle
    survived.  ``threads`` is accepted for compatibility; lanes are
    persistent (two waves per SIMD) and claim particles from a counter.
    """

    def __init__(self, line, flags=0x00, turns=1000, particles=1000, parameters=(), dp=0.,
                 context=None, options=None):
        self._init_common(flags, context, options)
        self.turns = int(turns)
        self.particles = int(particles)
        self.parameter_names = list(parameters)
        self.kernel = codegen.sa_kernel(line, flags, self.turns, self.parameter_names, dp)
        P = max(1, len(self.parameter_names))
        self._pool_host = host_buffer((self.particles, P), self.dtype)
        self.parameters = self._pool_host.numpy()
        self._lost_host = host_buffer((self.particles,), torch.int32)
        self.lost = self._lost_host.numpy().view(np.uint32)
        self._pool_dev = None
        self._lost_dev = None
        self._counter = None
        self.lanes = None

    def _buffers(self):
        if self._pool_dev is None:
            self._pool_dev = torch.empty(self._pool_host.shape, dtype=self.dtype, device=self.device)
            self._lost_dev = torch.empty((self.particles,), dtype=torch.int32, device=self.device)
            self._counter = torch.zeros((1,), dtype=torch.int32, device=self.device)

    @property
    def pool_device(self):
        self._buffers()
        return self._pool_dev

    @property
    def lost_device(self):
        self._buffers()
        return self._lost_dev

    # Persistent lanes per SIMD.  Two waves keep the FP64 pipe of a SIMD
    # busy (the 2^20-particle lossless scan runs as fast as with eight), and
    # with fewer lanes each lane refills from the counter more often, so the
    # end of a launch (lanes still tracking their last particle after the
    # counter runs out, beside lanes already idle) is a smaller share of it:
    # on a scan that loses 40 % of its particles early, 151 ms with two
    # waves per SIMD against 165 ms with eight, 216 ms with one lane per
    # particle; one wave per SIMD leaves the pipe waiting on latency.
    WAVES_PER_SIMD = 2

    # Order in which lanes claim particles: "amplitude" claims runs of
    # CLAIM_RUN consecutive particles by increasing summed normalised
    # amplitude (the runs likely to survive longest first), None in pool
    # order.  Claiming the long-lived particles first leaves the short-lived
    # ones for the end of the launch, where they fill the gaps of lanes still
    # on their last particle; whole runs keep a wave's refill reading and
    # writing consecutive entries.  Results are per particle and do not
    # depend on the order.
    claim_order = "amplitude"
    CLAIM_RUN = 64

    _AMPLITUDE_COLUMNS = ("x", "px", "y", "py", "dp")

    def order_device(self):
        """Claim order of the device pool (int32 permutation on the device),
        or None: the pool cut into runs of ``CLAIM_RUN`` consecutive
        particles, the full runs sorted by the sum over their particles of
        the sum over x, px, y, py, dp (the columns the pool has) of (value /
        largest |value| of that column)^2 (NaN values count as 0), a short
        last run claimed last.  Device work only: no host synchronisation."""
        if self.claim_order is None:
            return None
        if self.claim_order != "amplitude":
            raise ValueError(f"claim_order {self.claim_order!r}: 'amplitude' or None")
        R, n = self.CLAIM_RUN, self.particles
        runs = n // R
        cols = [self.parameter_names.index(c) for c in self._AMPLITUDE_COLUMNS if c in self.parameter_names]
        if runs < 2 or not cols:
            return None
        v = self._pool_dev[:runs * R, cols].nan_to_num(nan=0.0, posinf=0.0, neginf=0.0)
        s = v.abs().amax(dim=0)
        s = torch.where(s > 0, s, torch.ones_like(s))
        key = (v / s).square().view(runs, R * len(cols)).sum(dim=1)
        lane = torch.arange(R, dtype=torch.int32, device=v.device)
        order = (torch.argsort(key, stable=True).to(torch.int32)[:, None] * R + lane[None, :]).view(-1)
        if runs * R < n:
            order = torch.cat([order, torch.arange(runs * R, n, dtype=torch.int32, device=v.device)])
        return order

    def launch_config(self, lanes=None):
        block = self.kernel.block
        if lanes is None:
            lanes = self.resident_lanes(waves_per_simd=self.WAVES_PER_SIMD)
        lanes = max(1, min(lanes, self.particles))
        grid = (lanes + block - 1) // block
        return grid, block

    def launch(self, lanes=None):
        """Enqueue one pass on the current stream with the device pool as
        input (no host copies)."""
        self._buffers()
        if self.particles == 0:
            return
        if self.turns <= 0:
            self._lost_dev.zero_()
            return
        grid, block = self.launch_config(lanes)
        self.lanes = grid * block
        order = self.order_device()
        self._counter.zero_()
        self.module().launch(self.kernel.name, grid, block,
                             [self._pool_dev, self._lost_dev, self._counter, self.particles, grid * block,
                              order if order is not None else ctypes.c_void_p(0)])

    def run(self, threads=1):
        """Copy ``parameters`` to the device, track, copy ``lost`` back."""
        self._buffers()
        self._pool_dev.copy_(self._host_pool(), non_blocking=True)
        self.launch()
        self._lost_host.copy_(self._lost_dev, non_blocking=True)
        self.stream().synchronize()

    def executed_turns(self, lost=None):
        """Turns executed per particle: lost+1 for lost particles, turns for
        survivors."""
        lost = self.lost if lost is None else lost
        lost = np.asarray(lost, dtype=np.int64)
        return np.minimum(lost + 1, self.turns)
