"""Writes tests/golden/notebook_recon.npz from the reference prototype notebook.

swirl_lm (which holds Swirl-C's WENO and MUSCL face reconstruction) is not
vendored and cannot be imported.  The only in-repo text of these algorithms is
demo/swirl_c_prototype.ipynb, cell 4 (`weno5`, `muscl` and the stencil
helpers they call).  This script executes that cell's functions on seeded
inputs and stores inputs and outputs as a small fixture (data only), which
pins the oracle's `weno(k=3)` and `muscl` to the prototype.

Run in a container that has the reference at /root/reference:
  python tests/golden/make_notebook_vectors.py
"""

import json
import os

import numpy as np

NB = '/root/reference/demo/swirl_c_prototype.ipynb'
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'notebook_recon.npz')


def main():
  nb = json.load(open(NB))
  src = ''.join(nb['cells'][4]['source'])
  ns = {'np': np}
  exec(compile(src, 'swirl_c_prototype.ipynb[cell 4]', 'exec'), ns)  # pylint: disable=exec-used
  rng = np.random.default_rng(2024)
  cases = {}
  n = 24
  # Notebook layout: dim 'x' is axis 0.
  smooth = np.sin(np.linspace(0, 2 * np.pi, n))[:, None, None] * np.ones((n, 3, 2))
  step = np.where(np.arange(n) < n // 2, 1.0, 0.125)[:, None, None] * np.ones((n, 3, 2))
  noisy = rng.standard_normal((n, 3, 2))
  kinks = np.abs(np.linspace(-1, 1, n))[:, None, None] * np.ones((n, 3, 2))
  for name, v in (('smooth', smooth), ('step', step), ('noisy', noisy), ('kinks', kinks)):
    v = v.astype(np.float64)
    wn, wp = ns['weno5'](v.copy(), 'x')
    mn, mp = ns['muscl'](v.copy(), 'x')
    cases[f'{name}_in'] = v
    cases[f'{name}_weno5_neg'] = wn
    cases[f'{name}_weno5_pos'] = wp
    cases[f'{name}_muscl_neg'] = mn
    cases[f'{name}_muscl_pos'] = mp
  np.savez_compressed(OUT, **cases)
  print(OUT, sorted(cases))


if __name__ == '__main__':
  main()
