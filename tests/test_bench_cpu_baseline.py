"""The slab-parallel CPU oracle step that bench.py times as its CPU baseline
equals the serial oracle's RK3 step bit for bit (CPU only)."""

import multiprocessing as mp
from multiprocessing import shared_memory

import numpy as np

import bench
from oracle import swirlc_oracle as O


def test_parallel_oracle_step_equals_serial_step():
  ocfg, st = bench._kh_oracle_cfg(12)
  dt = np.float32(1e-3)
  ref = O.rk3(st, {}, ocfg, dt=dt)
  names = list(st)
  shape = (len(names),) + st[names[0]].shape
  nbytes = int(np.prod(shape)) * 4
  shm_in = shared_memory.SharedMemory(create=True, size=nbytes)
  shm_out = shared_memory.SharedMemory(create=True, size=nbytes)
  try:
    bench._SLAB.update({'in': shm_in.name, 'out': shm_out.name, 'shape': shape,
                        'h': ocfg['halo_width'], 'names': names, 'cfg': ocfg})
    slabs = [(0, 5), (5, 7), (7, 12)]
    with mp.get_context('fork').Pool(2) as pool:
      got = bench.parallel_rk3_step(pool, st, ocfg, dt, slabs, shm_in, shm_out, shape)
  finally:
    for m in (shm_in, shm_out):
      m.close()
      m.unlink()
  h = ocfg['halo_width']
  core = (slice(h, -h),) * 3
  for k in names:
    np.testing.assert_array_equal(got[k][core], ref[k][core], err_msg=k)
