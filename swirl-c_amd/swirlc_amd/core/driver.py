"""Host driver: cycles of time steps with state save and restore.

The reference's driver (core/driver.py:135-160) hands the time loop to
swirl_lm's driver.solver, which is not vendored with the reference.  The
contract visible through the reference's configuration is kept here on one
GPU:

  * the initial state comes from `customized_init_fn(replica_id, coordinates,
    cfg)`, or from `core.initializer.default_initializer` (all zeros,
    core/initializer.py:41-71, meant to be replaced by a loaded state);
  * the fields named in cfg.states_from_file are then loaded from the save
    file of step cfg.start_step;
  * cfg.num_cycles cycles of cfg.num_steps steps run from cfg.start_step, and
    the fields named in cfg.states_to_file are written after every cycle;
  * the configuration is written once as YAML (core/parameter.py:250-255).

Save files are numpy archives, one per saved step: `{prefix}-{step:010d}.npz`
holding one float32 array per field, shaped [nz+2h, nx+2h, ny+2h] (halo cells
included), plus `step` and `halo_width`; they load with
numpy.load(allow_pickle=False).
"""

import os

import numpy as np
import torch

from swirlc_amd.core import initializer
from swirlc_amd.core.solver import Solver


def save_path(prefix, step):
  return f'{prefix}-{int(step):010d}.npz'


def save_states(prefix, step, states, names, halo_width):
  """Writes the named fields of `states` (torch tensors or numpy arrays)."""
  d = os.path.dirname(prefix)
  if d:
    os.makedirs(d, exist_ok=True)
  arrays = {}
  for name in names:
    v = states[name]
    arrays[name] = (v.detach().cpu().numpy() if torch.is_tensor(v) else np.asarray(v)).astype(
        np.float32, copy=False)
  path = save_path(prefix, step)
  tmp = path + '.tmp.npz'
  np.savez(tmp, step=np.int64(step), halo_width=np.int64(halo_width), **arrays)
  os.replace(tmp, path)
  return path


def load_states(prefix, step, names):
  """Reads the named fields of the save file of `step` (numpy float32)."""
  path = save_path(prefix, step)
  with np.load(path, allow_pickle=False) as f:
    missing = [n for n in names if n not in f.files]
    if missing:
      raise KeyError(f'{path} lacks {missing}')
    if int(f['step']) != int(step):
      raise ValueError(f'{path} holds step {int(f["step"])}, expected {step}')
    return {n: np.array(f[n], dtype=np.float32) for n in names}


def init_state(customized_init_fn, cfg, device='cuda'):
  """The initial conservative fields as float32 tensors on `device`."""
  shape = (cfg.core_nz + 2 * cfg.halo_width, cfg.core_nx + 2 * cfg.halo_width,
           cfg.core_ny + 2 * cfg.halo_width)
  if customized_init_fn is None:
    customized_init_fn = initializer.default_initializer
  states = customized_init_fn(0, (0, 0, 0), cfg)
  out = {}
  for n in cfg.conservative_variable_names:
    v = states[n]
    v = v if torch.is_tensor(v) else torch.from_numpy(np.asarray(v, dtype=np.float32))
    if tuple(v.shape) != shape:
      raise ValueError(f'{n} has shape {tuple(v.shape)}, expected {shape}')
    out[n] = v.to(device=device, dtype=torch.float32).contiguous()
  return out


def solver(customized_init_fn, cfg, load_prefix=None, dump_prefix=None, cfl=None,
           use_graph=True, callback=None):
  """Runs cfg.num_cycles cycles of cfg.num_steps steps from cfg.start_step.

  Returns (state dict of the final step, final step index).  `callback(cycle,
  step, state)` runs after every cycle (after its save file is written).
  """
  state = init_state(customized_init_fn, cfg)
  step = int(cfg.start_step)
  from_file = list(cfg.states_from_file or [])
  if from_file:
    if load_prefix is None:
      raise ValueError('states_from_file is set but no load_prefix was given')
    for n, v in load_states(load_prefix, step, from_file).items():
      if tuple(v.shape) != tuple(state[n].shape):
        raise ValueError(f'{n} in the save file has shape {v.shape}')
      state[n].copy_(torch.from_numpy(v))
  to_file = list(cfg.states_to_file or [])
  if dump_prefix is not None:
    cfg.save_to_file(dump_prefix)
  sim = Solver(cfg, state, cfl=cfl, use_graph=use_graph)
  for cycle in range(int(cfg.num_cycles)):
    sim.step(int(cfg.num_steps))
    step += int(cfg.num_steps)
    if to_file and dump_prefix is not None:
      torch.cuda.synchronize()
      save_states(dump_prefix, step, sim.state, to_file, cfg.halo_width)
    if callback is not None:
      callback(cycle, step, sim.state)
  return sim.state, step
