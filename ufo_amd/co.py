"""ClosedOrbit (ufo/co.py:126-250) and the one-turn tangent map.

The reference finds each closed orbit with a per-particle Nelder-Mead search
on the squared one-turn residual (200 iterations by default, ~300 one-turn
passes).  Here each particle runs Newton's method on the one-turn map with
the Jacobian from the forward-tangent kernel: X <- X - (J - I)^-1 (F(X) - X),
a handful of one-turn passes to machine precision.  The objective is the
reference's: dp = 0, oodppo = 1, z = 0 (ufo/co.py:175-193), whatever the
flags say.
"""
from __future__ import annotations

import numpy as np
import torch

from . import tangent
from .sim import Simulation, host_buffer


class ClosedOrbit(Simulation):
    """Closed orbit (x, px, y, py) at the start of ``line`` for every
    particle's parameter set.  Same arguments and attributes as the reference
    (ufo/co.py:146-250): ``parameters`` [particles, max(1, P)], ``orbits``
    [particles, 4] after ``run()``.  ``iterations`` bounds the Newton steps;
    ``tol`` is the stopping residual max |F(X) - X| in metres/radians (in
    FP32, Newton also stops once a step no longer lowers the residual); an
    orbit that does not reach it is returned as the iterate with the
    smallest residual (the guess itself when its first turn is lost).
    ``steps`` holds the Newton steps each particle took and ``converged``
    whether its residual reached ``tol``."""

    def __init__(self, line, flags=0x00, particles=1000, parameters=(), dp=0., iterations=200,
                 context=None, options=None, tol=1e-15):
        self._init_common(flags, context, options)
        self.particles = int(particles)
        self.parameter_names = list(parameters)
        self.iterations = int(iterations)
        self.kernel = tangent.co_kernel(line, flags, self.parameter_names, min(self.iterations, 64), tol)
        P = max(1, len(self.parameter_names))
        self._pool_host = host_buffer((self.particles, P), self.dtype)
        self.parameters = self._pool_host.numpy()
        self._orbits_host = host_buffer((self.particles, 4), self.dtype)
        self.orbits = self._orbits_host.numpy()
        self._steps_host = host_buffer((self.particles,), torch.int32)
        self.steps = self._steps_host.numpy()
        self._conv_host = host_buffer((self.particles,), torch.int32)
        self._dev = None

    @property
    def converged(self):
        """bool [particles]: the residual reached ``tol`` (after ``run()``)."""
        return self._conv_host.numpy() != 0

    def run(self, threads=1):
        if self._dev is None:
            self._dev = (torch.empty(self._pool_host.shape, dtype=self.dtype, device=self.device),
                         torch.empty(self._orbits_host.shape, dtype=self.dtype, device=self.device),
                         torch.empty(self._steps_host.shape, dtype=torch.int32, device=self.device),
                         torch.empty(self._conv_host.shape, dtype=torch.int32, device=self.device))
        pool, orbits, steps, conv = self._dev
        pool.copy_(self._host_pool(), non_blocking=True)
        if self.particles:
            block = self.kernel.block
            grid = (self.particles + block - 1) // block
            self.module().launch(self.kernel.name, grid, block, [pool, orbits, steps, conv, self.particles])
        self._orbits_host.copy_(orbits, non_blocking=True)
        self._steps_host.copy_(steps, non_blocking=True)
        self._conv_host.copy_(conv, non_blocking=True)
        self.stream().synchronize()


class TangentMap(Simulation):
    """State and transverse Jacobian after ``turns`` turns, per particle.

    ``parameters``/``dp``/flags as for Track.  After ``run()``: ``state``
    [particles, 4] (x, px, y, py) and ``jacobian`` [particles, 4, 4] with
    jacobian[p, i, j] = d state_i / d initial coordinate j (dp held fixed)."""

    def __init__(self, line, flags=0x00, turns=1, particles=1, parameters=(), dp=0., context=None,
                 options=None):
        self._init_common(flags, context, options)
        self.turns = int(turns)
        self.particles = int(particles)
        self.parameter_names = list(parameters)
        self.kernel = tangent.jvp_kernel(line, flags, self.turns, self.parameter_names, dp)
        P = max(1, len(self.parameter_names))
        self._pool_host = host_buffer((self.particles, P), self.dtype)
        self.parameters = self._pool_host.numpy()
        self._out_host = host_buffer((self.particles, 20), self.dtype)
        self._dev = None

    @property
    def state(self):
        return self._out_host.numpy()[:, :4]

    @property
    def jacobian(self):
        cols = self._out_host.numpy()[:, 4:].reshape(self.particles, 4, 4)   # [p, column j, row i]
        return np.transpose(cols, (0, 2, 1))

    def run(self, threads=1):
        if self._dev is None:
            self._dev = (torch.empty(self._pool_host.shape, dtype=self.dtype, device=self.device),
                         torch.empty(self._out_host.shape, dtype=self.dtype, device=self.device))
        pool, out = self._dev
        pool.copy_(self._host_pool(), non_blocking=True)
        if self.particles:
            block = self.kernel.block
            grid = (self.particles + block - 1) // block
            self.module().launch(self.kernel.name, grid, block, [pool, out, self.particles])
        self._out_host.copy_(out, non_blocking=True)
        self.stream().synchronize()


def symplectic_error(J):
    """max |J^T S J - S| for 4x4 transverse Jacobians J [..., 4, 4]."""
    S = np.array([[0, 1, 0, 0], [-1, 0, 0, 0], [0, 0, 0, 1], [0, 0, -1, 0]], dtype=np.float64)
    J = np.asarray(J, dtype=np.float64)
    return np.max(np.abs(np.swapaxes(J, -1, -2) @ S @ J - S), axis=(-2, -1))
