"""ufo_launch_noalloc, the entry point that allocates nothing, against
ufo_launch on the aperture kernel (its scratch buffer is the work-pool
counter, ufo/sa.py:122): identical loss turns on two calls in a row, the
second on a smaller grid, with fewer lanes than particles so that every lane
refills from the counter the library clears."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _aperture(ring, flags, turns, grid):
    import ufo_amd as ufo
    sa = ufo.StableAperture(ring, flags=flags, turns=turns, particles=len(grid), parameters=["x", "y"])
    sa.parameters[:] = grid
    sa.run()                                     # ufo_launch with the driver's own counter
    return sa


def _noalloc(sa, lanes):
    """The same pass through ufo_launch_noalloc into a fresh lost buffer
    filled with a sentinel (an entry the kernel skips stays 0xffffffff)."""
    lost = torch.full((sa.particles,), -1, dtype=torch.int32, device=sa.device)
    grid, block = sa.launch_config(lanes)
    print(f"[no_alloc] particles={sa.particles} lanes={grid * block}", flush=True)
    sa.module().launch_noalloc(sa.kernel.name, grid, block,
                               [sa.pool_device, lost, None, sa.particles, grid * block, None],
                               scratch_arg=2, scratch_bytes=4)
    torch.cuda.synchronize()
    return lost.cpu().numpy().view(np.uint32)


def test_noalloc_launch_matches_launch_on_two_calls():
    import ufo_amd as ufo
    from ufo_amd import lattices
    flags = ufo.KICK | ufo.DOUBLE_PRECISION | ufo.FIVED
    ring = lattices.dba_ring(16)
    turns = 256
    big = lattices.da_grid(128, 64, xlim=0.03, ymax=0.02)        # 8192 points, part of them lost
    small = lattices.da_grid(50, 40, xlim=0.025, ymax=0.015)     # 2000 points
    sa1 = _aperture(ring, flags, turns, big)
    got1 = _noalloc(sa1, lanes=1024)
    assert np.array_equal(got1, sa1.lost), "first call differs from ufo_launch"
    n_lost = int(np.sum(sa1.lost < turns))
    assert 0 < n_lost < len(big), n_lost
    sa2 = _aperture(ring, flags, turns, small)
    assert sa2.module() is sa1.module()                         # same kernel, the array keeps call 1's count
    got2 = _noalloc(sa2, lanes=256)
    assert np.array_equal(got2, sa2.lost), "second (smaller) call differs from ufo_launch"
    print(f"[no_alloc] call 1: {len(big)} particles, {n_lost} lost; call 2: {len(small)} particles, "
          f"{int(np.sum(sa2.lost < turns))} lost; loss turns identical", flush=True)


def test_noalloc_rejects_oversized_scratch():
    import ufo_amd as ufo
    from ufo_amd import lattices, runtime
    sa = ufo.StableAperture(lattices.fodo_ring(2), flags=ufo.KICK | ufo.DOUBLE_PRECISION | ufo.FIVED, turns=4,
                            particles=64, parameters=["x", "y"])
    lost = torch.zeros((64,), dtype=torch.int32, device=sa.device)
    with pytest.raises(runtime.UfoError, match="scratch of 1048577 bytes"):
        sa.module().launch_noalloc(sa.kernel.name, 1, 64, [sa.pool_device, lost, None, 64, 64, None],
                                   scratch_arg=2, scratch_bytes=(1 << 20) + 1)
    with pytest.raises(runtime.UfoError, match="outside the 6 kernel arguments"):
        sa.module().launch_noalloc(sa.kernel.name, 1, 64, [sa.pool_device, lost, None, 64, 64, None],
                                   scratch_arg=6, scratch_bytes=4)
