"""Track: tracking with the particle coordinates sampled at observation
points every turn (ufo/track.py:19-139)."""
from __future__ import annotations

import torch

from . import codegen
from .sim import Simulation, host_buffer


class Track(Simulation):
    """Track a bunch through ``line`` for ``turns`` turns, recording
    (x, px, y, py, z, dp) at the positions ``where`` (element index plus
    fraction of that element; -1 = end of line) every turn.

    Same arguments and attributes as the reference (ufo/track.py:40-139):
    ``parameters`` is a host array [particles, len(parameters)] the user fills
    before ``run()``; ``tracks`` is [particles, len(where)*turns, 6], samples
    in lattice order, turn after turn.  Samples after a particle is lost stay
    NaN (the reference leaves them unwritten); under FIVED, z and dp hold 0
    and the ``dp`` argument.  ``run()`` blocks until ``tracks`` is filled.
    ``threads`` is accepted for compatibility; the launch is sized from the
    particle count (one lane per particle).
    """

    def __init__(self, line, flags=0x00, turns=1000, particles=1000, parameters=(), where=(), dp=0.,
                 context=None, options=None):
        self._init_common(flags, context, options)
        where = list(where) if isinstance(where, (list, tuple)) else [where]
        self.turns = int(turns)
        self.particles = int(particles)
        self.parameter_names = list(parameters)
        self.count = len(where)
        total = self.count * self.turns
        if total < 1:
            raise Exception("No valid observation points found in 'where'")
        self.kernel = codegen.track_kernel(line, flags, self.turns, self.parameter_names, where, dp)
        if self.kernel.dumps_per_turn > self.count:
            raise Exception("more observation points than entries in 'where'")
        flat = line.flatten()
        self.where = [(idx, flat[int(idx)]) for idx in where]
        self.samples = total
        P = max(1, len(self.parameter_names))
        self._pool_host = host_buffer((self.particles, P), self.dtype)
        self.parameters = self._pool_host.numpy()
        self._tracks_host = host_buffer((total, 6, self.particles), self.dtype)
        self.tracks = self._tracks_host.permute(2, 0, 1).numpy()
        self._pool_dev = None
        self._tracks_dev = None
        self._state_dev = None
        self._run_stream = None
        self._copy_stream = None
        self._events = []
        # run() splits the turns over this many launches and copies each
        # finished slab of samples back while the next launch runs; None
        # picks 16 when the samples exceed 16 MB (else one launch).
        self.copy_chunks = None

    def _buffers(self):
        if self._pool_dev is None:
            self._pool_dev = torch.empty(self._pool_host.shape, dtype=self.dtype, device=self.device)
            self._tracks_dev = torch.empty(self._tracks_host.shape, dtype=self.dtype, device=self.device)
            self._state_dev = torch.empty((4, max(1, self.particles)), dtype=self.dtype, device=self.device)

    @property
    def tracks_device(self):
        """Device tensor [particles, samples, 6] (a strided view of the
        [sample][coordinate][particle] buffer) after ``run()``."""
        self._buffers()
        return self._tracks_dev.permute(2, 0, 1)

    def _launch_turns(self, turn0, turn1):
        block = self.kernel.block
        grid = (self.particles + block - 1) // block
        self.module().launch(self.kernel.name, grid, block,
                             [self._pool_dev, self._tracks_dev, self._state_dev, self.particles, turn0, turn1])

    def _fill_unwritten(self):
        """NaN into the samples no launch writes: observation points past
        turns * dumps_per_turn (fractional points on thin elements, never
        observed, as in the reference).  The kernel writes every other
        sample, NaN after a particle's loss."""
        first = self.turns * self.kernel.dumps_per_turn
        if first < self._tracks_dev.shape[0]:
            self._tracks_dev[first:].fill_(float("nan"))

    def launch(self):
        """Enqueue one tracking pass on the current stream with the device
        pool as input (no host copies)."""
        self._buffers()
        self._fill_unwritten()
        if self.particles == 0:
            return
        self._launch_turns(0, self.turns)

    def _chunks(self):
        n = self.copy_chunks
        if n is None:
            n = 16 if self._tracks_dev.numel() * self._tracks_dev.element_size() > (16 << 20) else 1
        return max(1, min(int(n), self.turns))

    def run(self, threads=1):
        """Copy ``parameters`` to the device, track, copy ``tracks`` back.
        Split runs (``copy_chunks`` launches) go to a stream of this object
        and a second one copies each finished slab of samples to the host
        while the next launch runs.  The two streams are created one after
        the other, which puts them on different hardware queues (a copy on
        the queue of the launches would wait for them)."""
        self._buffers()
        caller = self.stream()
        n = self._chunks()
        if self.particles == 0 or n == 1:
            self._pool_dev.copy_(self._host_pool(), non_blocking=True)
            self._fill_unwritten()
            if self.particles:
                self._launch_turns(0, self.turns)
            self._tracks_host.copy_(self._tracks_dev, non_blocking=True)
            caller.synchronize()
            return
        if self._copy_stream is None:
            self._run_stream = torch.cuda.Stream(self.device)
            self._copy_stream = torch.cuda.Stream(self.device)
            self._events = []
        runner, copier = self._run_stream, self._copy_stream
        while len(self._events) < n + 1:
            self._events.append(torch.cuda.Event())
        self._events[n].record(caller)           # work the caller queued on these buffers
        runner.wait_event(self._events[n])
        bounds = [self.turns * k // n for k in range(n + 1)]
        d = self.kernel.dumps_per_turn
        with torch.cuda.stream(runner):
            self._pool_dev.copy_(self._host_pool(), non_blocking=True)
            self._fill_unwritten()
            for k, (t0, t1) in enumerate(zip(bounds[:-1], bounds[1:])):
                self._launch_turns(t0, t1)
                done = self._events[k]
                done.record(runner)
                copier.wait_event(done)
                # samples are numbered turn * dumps_per_turn + k; entries past
                # turns * dumps_per_turn (points that are never written) go
                # with the last slab
                a, b = t0 * d, (t1 * d if t1 < self.turns else self.samples)
                with torch.cuda.stream(copier):
                    self._tracks_host[a:b].copy_(self._tracks_dev[a:b], non_blocking=True)
        copier.synchronize()
        runner.synchronize()
