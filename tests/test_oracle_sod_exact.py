"""The fp64 oracle's Sod shock tube (tests/golden/sod_fp64.npz) against the
exact Riemann solution at t = 0.2: pins the whole oracle step (WENO5, HLLC,
SSP-RK3, Neumann halos) by the mathematics of the problem."""

import os

import numpy as np

GAMMA = 1.4
HERE = os.path.dirname(os.path.abspath(__file__))


def _exact_sod(x, t, left=(1.0, 0.0, 1.0), right=(0.125, 0.0, 0.1), x0=0.5):
  """Exact solution of the Riemann problem (rho, u, p) for an ideal gas
  (Toro, Riemann Solvers and Numerical Methods, ch. 4): Newton iteration on
  the star pressure, then sampling of the wave pattern at x / t."""
  g = GAMMA
  rl, ul, pl = left
  rr, ur, pr = right
  cl, cr = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)

  def f(p, rk, pk, ck):
    if p > pk:  # shock
      a, b = 2.0 / ((g + 1) * rk), (g - 1) / (g + 1) * pk
      q = np.sqrt(a / (p + b))
      return (p - pk) * q, q * (1.0 - 0.5 * (p - pk) / (p + b))
    r = (p / pk) ** ((g - 1) / (2 * g))  # rarefaction
    return 2 * ck / (g - 1) * (r - 1.0), (p / pk) ** (-(g + 1) / (2 * g)) / (rk * ck)

  p = 0.5 * (pl + pr)
  for _ in range(100):
    fl, dl = f(p, rl, pl, cl)
    fr, dr = f(p, rr, pr, cr)
    dp = (fl + fr + ur - ul) / (dl + dr)
    p = max(p - dp, 1e-12)
    if abs(dp) < 1e-14 * p:
      break
  fl, _ = f(p, rl, pl, cl)
  fr, _ = f(p, rr, pr, cr)
  us = 0.5 * (ul + ur) + 0.5 * (fr - fl)
  s = (np.asarray(x) - x0) / t
  rho = np.empty_like(s)
  u = np.empty_like(s)
  pp = np.empty_like(s)
  # Left rarefaction (p* < pl) and right shock (p* > pr), as in Sod's problem.
  assert p < pl and p > pr
  rsl = rl * (p / pl) ** (1.0 / g)
  csl = cl * (p / pl) ** ((g - 1) / (2 * g))
  head, tail = ul - cl, us - csl
  rsr = rr * ((p / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * p / pr + 1.0))
  sshock = ur + cr * np.sqrt((g + 1) / (2 * g) * p / pr + (g - 1) / (2 * g))
  for i, si in enumerate(s):
    if si < head:
      rho[i], u[i], pp[i] = rl, ul, pl
    elif si < tail:
      ui = 2.0 / (g + 1) * (cl + (g - 1) / 2 * ul + si)
      ci = 2.0 / (g + 1) * (cl + (g - 1) / 2 * (ul - si))
      rho[i] = rl * (ci / cl) ** (2.0 / (g - 1))
      u[i], pp[i] = ui, pl * (ci / cl) ** (2 * g / (g - 1))
    elif si < us:
      rho[i], u[i], pp[i] = rsl, us, p
    elif si < sshock:
      rho[i], u[i], pp[i] = rsr, us, p
    else:
      rho[i], u[i], pp[i] = rr, ur, pr
  return rho, u, pp, (p, us, rsl, rsr, sshock)


def test_exact_solver_textbook_star_state():
  # Toro, table 4.2 (test 1): p* = 0.30313, u* = 0.92745.
  _, _, _, (p, us, rsl, rsr, _) = _exact_sod(np.array([0.5]), 0.2)
  assert abs(p - 0.30313) < 1e-5 and abs(us - 0.92745) < 1e-5
  assert abs(rsl - 0.42632) < 1e-5 and abs(rsr - 0.26557) < 1e-5


def test_oracle_sod_fp64_matches_exact_solution():
  ref = np.load(os.path.join(HERE, 'golden', 'sod_fp64.npz'))
  n = int(ref['n'])
  t = float(ref['dt']) * int(ref['steps'])
  assert abs(t - 0.2) < 1e-12
  x = (np.arange(n) + 0.5) / n
  rho, u, p, _ = _exact_sod(x, t)
  # WENO5 at 512 cells: L1 errors of a few 1e-3 (smeared shock and contact).
  for got, ex, tol in ((ref['rho'], rho, 3e-3), (ref['u'], u, 5e-3), (ref['p'], p, 3e-3)):
    assert np.mean(np.abs(got - ex)) < tol
  # Away from the discontinuities and the kinks at the rarefaction ends the
  # solution is close to exact (max error a few 1e-3 within 0.03 of them).
  feats = (0.5 - 1.18322 * t, 0.5 + (0.92745 - 0.99739) * t, 0.5 + 0.92745 * t,
           0.5 + 1.75216 * t)
  far = np.ones(n, bool)
  for xf in feats:
    far &= np.abs(x - xf) > 0.03
  assert far.sum() > 350
  for got, ex, tol in ((ref['rho'], rho, 2.5e-3), (ref['u'], u, 5e-3), (ref['p'], p, 2.5e-3)):
    assert np.max(np.abs(got[far] - ex[far])) < tol
