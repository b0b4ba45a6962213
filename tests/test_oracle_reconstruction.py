"""Pins the oracle's face reconstruction (WENO-5, WENO-3, MUSCL).

swirl_lm's `interpolation.weno` / `flux_limiter` are not vendored, so no
reference test pins them (every reference RHS test mocks them,
equations/compressible_navier_stokes_test.py:36-59).  The pins are:
  * golden vectors from the in-repo prototype notebook
    (demo/swirl_c_prototype.ipynb cell 4, tests/golden/notebook_recon.npz),
    compared after moving from the notebook's i+1/2 storage to swirl_lm's
    i-1/2 storage;
  * mathematical invariants: WENO-5 reproduces face values of quadratics from
    their cell averages exactly, WENO-3 of linear functions; both keep the
    two sides of a step (ENO); both reach their design order (5, 3) on
    smooth monotone data, which pins the linear weights; MUSCL creates no new
    extrema (TVD).
"""

import os

import numpy as np
import pytest

from oracle import swirlc_oracle as O

GOLD = os.path.join(os.path.dirname(__file__), 'golden', 'notebook_recon.npz')


@pytest.fixture(scope='module')
def gold():
  return np.load(GOLD)


@pytest.mark.parametrize('case', ['smooth', 'step', 'noisy', 'kinks'])
def test_weno5_matches_notebook(gold, case):
  v = gold[f'{case}_in']
  n = v.shape[0]
  neg, pos = O.weno(v, 'z', 3)  # notebook axis 0 is the oracle's z axis
  nb_neg, nb_pos = gold[f'{case}_weno5_neg'], gold[f'{case}_weno5_pos']
  i = np.arange(2, n - 2)     # notebook face i+1/2 from cell i (neg)
  np.testing.assert_allclose(neg[i + 1], nb_neg[i], rtol=1e-13, atol=1e-13)
  i = np.arange(1, n - 3)     # notebook face i+1/2 from cell i+1 (pos)
  np.testing.assert_allclose(pos[i + 1], nb_pos[i], rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize('case', ['smooth', 'step', 'noisy', 'kinks'])
def test_muscl_matches_notebook(gold, case):
  v = gold[f'{case}_in']
  n = v.shape[0]
  neg, pos = O.muscl(v, 'z')
  nb_neg, nb_pos = gold[f'{case}_muscl_neg'], gold[f'{case}_muscl_pos']
  i = np.arange(0, n - 1)
  np.testing.assert_allclose(neg[i + 1], nb_neg[i], rtol=1e-13, atol=1e-13)
  np.testing.assert_allclose(pos[i + 1], nb_pos[i], rtol=1e-13, atol=1e-13)


def _cell_avg_poly(coef, n, h=0.1):
  """Cell averages and face values of p(x) = sum coef[k] x^k on n cells."""
  edges = np.arange(n + 1) * h
  P = np.polynomial.polynomial.Polynomial(coef)
  I = P.integ()
  avg = (I(edges[1:]) - I(edges[:-1])) / h
  return avg, P(edges)


@pytest.mark.parametrize('dim', ['x', 'y', 'z'])
def test_weno5_exact_on_quadratics(dim):
  n = 20
  avg, faces = _cell_avg_poly([0.3, -1.2, 2.5], n)
  shape = [4, 4, 4]
  ax = O.AXIS[dim]
  shape[ax] = n
  bshape = [1, 1, 1]
  bshape[ax] = n
  v = np.broadcast_to(avg.reshape(bshape), shape).copy()
  neg, pos = O.weno(v, dim, 3)
  take = lambda a: np.moveaxis(a, ax, 0)[:, 0, 0]
  # Face i-1/2 is the left edge of cell i: faces[i].
  np.testing.assert_allclose(take(neg)[3:n - 2], faces[3:n - 2], rtol=0, atol=1e-10)
  np.testing.assert_allclose(take(pos)[2:n - 2], faces[2:n - 2], rtol=0, atol=1e-10)


def test_weno3_exact_on_linear():
  n = 16
  avg, faces = _cell_avg_poly([0.7, -2.0], n)
  v = np.broadcast_to(avg[None, :, None], (3, n, 3)).copy()
  neg, pos = O.weno(v, 'x', 2)
  np.testing.assert_allclose(neg[0, 2:n - 1, 0], faces[2:n - 1], atol=1e-12)
  np.testing.assert_allclose(pos[0, 1:n - 1, 0], faces[1:n - 1], atol=1e-12)


@pytest.mark.parametrize('k', [2, 3])
def test_weno_keeps_step_sides(k):
  n = 20
  v = np.where(np.arange(n) < n // 2, 1.0, 0.125)[None, :, None] * np.ones((2, n, 2))
  neg, pos = O.weno(v, 'x', k)
  j = n // 2  # face between cells j-1 (1.0) and j (0.125)
  assert abs(neg[0, j, 0] - 1.0) < 1e-6
  assert abs(pos[0, j, 0] - 0.125) < 1e-6


def test_muscl_is_tvd_on_step_and_noise():
  rng = np.random.default_rng(1)
  for v1 in (np.where(np.arange(30) < 15, 1.0, 0.125), rng.standard_normal(30)):
    v = v1[None, :, None] * np.ones((2, 30, 2))
    neg, pos = O.muscl(v, 'x')
    for i in range(2, 28):
      lo, hi = min(v1[i - 1], v1[i]), max(v1[i - 1], v1[i])
      assert lo - 1e-12 <= neg[0, i, 0] <= hi + 1e-12
      assert lo - 1e-12 <= pos[0, i, 0] <= hi + 1e-12


def test_weno_fp32_tracks_fp64():
  rng = np.random.default_rng(7)
  v = (1.0 + 0.3 * rng.standard_normal((6, 40, 5)))
  n64, p64 = O.weno(v, 'x', 3)
  n32, p32 = O.weno(v.astype(np.float32), 'x', 3)
  assert n32.dtype == np.float32
  np.testing.assert_allclose(n32, n64, rtol=2e-5, atol=2e-5)
  np.testing.assert_allclose(p32, p64, rtol=2e-5, atol=2e-5)


@pytest.mark.parametrize('k,order', [(3, 4.5), (2, 2.7)])
def test_weno_convergence_order_on_smooth_monotone_data(k, order):
  """WENO-JS reaches its design order (5 for k=3, 3 for k=2) on cell averages
  of a smooth function with no critical point.  Exactness on polynomials
  holds for any convex weights; the order pins the linear weights d (with
  the wrong d, WENO-5 drops to 3rd and WENO-3 to 2nd order)."""
  errs = []
  for n in (40, 80, 160):
    h = 1.0 / n
    edges = np.arange(n + 1) * h
    avg = (np.exp(edges[1:]) - np.exp(edges[:-1])) / h
    v = np.broadcast_to(avg[None, :, None], (2, n, 2)).copy()
    neg, pos = O.weno(v, 'x', k)
    # Face i-1/2 (left edge of cell i) stored at i; skip the stencil's reach.
    faces = np.exp(edges[:-1])
    sl = slice(k + 1, n - k)
    errs.append(max(np.abs(neg[0, sl, 0] - faces[sl]).max(),
                    np.abs(pos[0, sl, 0] - faces[sl]).max()))
  rates = [np.log2(errs[i] / errs[i + 1]) for i in range(2)]
  assert min(rates) >= order, (errs, rates)
