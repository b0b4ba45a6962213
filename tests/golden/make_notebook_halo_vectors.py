"""Writes tests/golden/notebook_halo.npz from the reference prototype notebook.

swirl_lm's halo exchange (called by boundary/boundary.py:77-86) is not
vendored.  The in-repo statement of its semantics is
demo/swirl_c_prototype.ipynb cell 2 (`update_boundary`, which applies
`periodic` / `dirichlet` / `neumann` dimension after dimension over whole
planes).  The reference's own boundary tests check only face-interior halo
cells, so the values that edge and corner halo cells take (cells that lie in
the halo of two or three dimensions, written once per dimension) are fixed by
that sequential fill alone.  This script executes the cell's functions on
seeded arrays with garbage in every halo cell and stores inputs, BCs and
outputs as a small fixture (data only).

The notebook indexes arrays [x][y][z]; the fixture is stored in the
[z][x][y] layout of swirl-c's FlowFieldVar (common/types.py:31-33).

Run in a container that has the reference at /root/reference:
  python tests/golden/make_notebook_halo_vectors.py
"""

import enum
import json
import os

import numpy as np

NB = '/root/reference/demo/swirl_c_prototype.ipynb'
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'notebook_halo.npz')

# BC codes stored in the fixture: 0 periodic, 1 Dirichlet scalar, 2 Neumann,
# 3 Dirichlet list of per-layer scalars, 4 Dirichlet list of per-layer planes.
CASES = {
    # name: (n = (nx, ny, nz), h, {dim: ((code, value), (code, value))})
    'dirichlet_neumann_mix': ((6, 5, 4), 3, {
        'x': ((1, 8.0), (2, 0.5)),
        'y': ((0, 0.0), (0, 0.0)),
        'z': ((2, -1.0), (3, [1.0, 2.0, 3.0]))}),
    'all_walls': ((5, 4, 6), 3, {
        'x': ((2, 0.25), (2, -0.75)),
        'y': ((1, -2.0), (1, 4.0)),
        'z': ((2, 1.5), (1, 9.0))}),
    'planes_and_periodic_z': ((4, 6, 5), 2, {
        'x': ((4, None), (2, 0.125)),
        'y': ((2, -0.5), (3, [7.0, 11.0])),
        'z': ((0, 0.0), (0, 0.0))}),
    'neumann_everywhere': ((3, 3, 3), 4, {
        'x': ((2, 1.0), (2, 2.0)),
        'y': ((2, 3.0), (2, 4.0)),
        'z': ((2, 5.0), (2, 6.0))}),
}


def to_zxy(v):
  return np.ascontiguousarray(np.transpose(v, (2, 0, 1)))


def main():
  nb = json.load(open(NB))
  src = ''.join(nb['cells'][2]['source'])
  ns = {'np': np, 'enum': enum}
  exec(compile(src, 'swirl_c_prototype.ipynb[cell 2]', 'exec'), ns)  # pylint: disable=exec-used
  bct = ns['BoundaryCondition']
  rng = np.random.default_rng(77)
  out = {}
  for name, (n, h, spec) in CASES.items():
    nx, ny, nz = n
    shape_xyz = (nx + 2 * h, ny + 2 * h, nz + 2 * h)
    v = rng.standard_normal(shape_xyz) * 10.0  # garbage in every halo cell
    v[h:-h, h:-h, h:-h] = rng.uniform(-1.0, 1.0, (nx, ny, nz))
    bc = {}
    codes = np.zeros((3, 2), np.int32)
    values = np.zeros((3, 2, h), np.float64)
    for d, dim in enumerate('xyz'):
      bc[dim] = {}
      for face in (0, 1):
        code, val = spec[dim][face]
        codes[d, face] = code
        if code == 0:
          bc[dim][face] = (bct.PERIODIC, None)
        elif code == 1:
          bc[dim][face] = (bct.DIRICHLET, val)
          values[d, face, :] = val
        elif code == 2:
          bc[dim][face] = (bct.NEUMANN, val)
          values[d, face, :] = val
        elif code == 3:
          bc[dim][face] = (bct.DIRICHLET, list(val))
          values[d, face, :] = val
        else:
          # Per-layer planes spanning the other two axes ([y][z] for x).
          other = [s for k, s in enumerate(shape_xyz) if k != d]
          planes = [rng.uniform(-3.0, 3.0, other) for _ in range(h)]
          bc[dim][face] = (bct.DIRICHLET, planes)
          # Stored in [z][x][y] order of the two remaining axes.
          key = f'{name}_plane_{dim}{face}'
          if dim == 'x':  # plane [y][z] -> [z][y]
            out[key] = np.stack([p.T for p in planes])
          elif dim == 'y':  # plane [x][z] -> [z][x]
            out[key] = np.stack([p.T for p in planes])
          else:  # plane [x][y] -> [x][y]
            out[key] = np.stack(planes)
    res = ns['update_boundary']({'v': v.copy()}, {'halo_width': h, 'bc': {'v': bc}})['v']
    out[f'{name}_in'] = to_zxy(v)
    out[f'{name}_out'] = to_zxy(res)
    out[f'{name}_codes'] = codes
    out[f'{name}_values'] = values
    out[f'{name}_halo'] = np.int32(h)
  np.savez_compressed(OUT, **out)
  print(OUT, sorted(out))


if __name__ == '__main__':
  main()
