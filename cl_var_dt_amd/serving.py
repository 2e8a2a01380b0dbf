"""Serving: continuous batching of Langevin runs on one GPU, and the secant
mu search of the reference's orchestration as a client of it.

`Server` wraps clv_server_* (include/clv.h): runs are submitted at any time,
admitted into device-resident slot pools by HBM budget while others are in
flight, and return their observables and time-to-result.  `secant_method`
restates the search loop of Secant_method_function (mylib.py:599-706) with its
new_mu / check_to_stop / change_guess (mylib.py:549-593), each density a mean
over independent server runs with its standard error; `secant_mu_search` is
the bare secant iteration on the same new_mu; `windowed_means` restates the
burn-in-cut running means of mylib.py:384-404.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Callable, Sequence

from . import _abi
from ._abi import Params, Request, Result, ServerInfo, Shape, lib


@dataclass
class RunResult:
    id: int
    density: float
    n_tot: float
    Sz: float
    energy: float
    count: int
    steps: int
    t_langevin: float
    latency_s: float
    queue_s: float
    raw: Result


class Server:
    def __init__(self, precision: int = _abi.CLV_FP32, hbm_budget_bytes: int = 0, device: int = 0):
        h = C.c_void_p()
        st = lib().clv_server_create(device, precision, int(hbm_budget_bytes), C.byref(h))
        if st != _abi.CLV_OK:
            raise _abi.ClvError(st, "clv_server_create failed")
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().clv_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def clock() -> float:
        return float(lib().clv_server_clock())

    def submit(self, shape: Shape, params: Params, nsteps: int, eval_every: int = 10,
               eval_from: int = 0, n0: float = 1.0, var: float = 0.1) -> int:
        req = Request(shape=shape, params=params, nsteps=nsteps, eval_every=eval_every,
                      reserved=0, eval_from=eval_from, n0=n0, var=var)
        rid = C.c_int64()
        st = lib().clv_server_submit(self._h, C.byref(req), C.byref(rid))
        if st != _abi.CLV_OK:
            raise _abi.ClvError(st, lib().clv_server_error(self._h).decode())
        return rid.value

    def cancel(self, rid: int) -> None:
        """Cancel a queued or running run (clv_server_cancel); its wait then
        raises ClvError with status CLV_ERR_CANCELLED."""
        st = lib().clv_server_cancel(self._h, rid)
        if st != _abi.CLV_OK:
            raise _abi.ClvError(st, lib().clv_server_error(self._h).decode())

    def poll(self, rid: int) -> Result:
        r = Result()
        _abi.check(lib().clv_server_poll(self._h, rid, C.byref(r)))
        return r

    def release(self, rid: int) -> None:
        """Forget a finished run's record (clv_server_release)."""
        st = lib().clv_server_release(self._h, rid)
        if st != _abi.CLV_OK:
            raise _abi.ClvError(st, lib().clv_server_error(self._h).decode())

    def wait(self, rid: int, timeout: float = 600.0, release: bool = False) -> RunResult:
        """Block until run `rid` finishes; release=True also forgets its record."""
        r = Result()
        st = lib().clv_server_wait(self._h, rid, timeout, C.byref(r))
        if st != _abi.CLV_OK:
            raise _abi.ClvError(st, f"run {rid} not done after {timeout} s")
        if release:
            self.release(rid)
        if r.status < 0:
            raise _abi.ClvError(r.status, lib().clv_server_error(self._h).decode())
        return RunResult(id=r.id, density=r.obs.density, n_tot=r.obs.n_tot, Sz=r.obs.Sz,
                         energy=r.obs.energy, count=r.obs.count, steps=r.stats.steps,
                         t_langevin=r.stats.t_langevin, latency_s=r.t_done - r.t_submit,
                         queue_s=r.t_admit - r.t_submit, raw=r)

    def info(self) -> ServerInfo:
        i = ServerInfo()
        _abi.check(lib().clv_server_info_get(self._h, C.byref(i)))
        return i


class JError(ValueError):
    """No earlier point left for a secant step after the first (mylib.py:549-553)."""


class GuessError(ValueError):
    """The first secant step from the two initial guesses failed (mylib.py:555-558)."""


def new_mu(r: Sequence[float], mu: Sequence[float], i: int | None = None) -> float:
    """Secant step of new_mu (mylib.py:561-580) on residuals r = observable - target:
        mu_new = (r[i+1] mu[j] - r[j] mu[i+1]) / (r[i+1] - r[j]),  j = i, i-1, ...
    i defaults to the last pair (i = len(r) - 2).  A step is refused and j moves
    to an older point when mu_new lies in [10 max(mu), 0] -- the reference's
    guard, kept as written: it can only fire when no mu so far is positive --
    or, unlike the reference, when r[i+1] == r[j] (there numpy returns inf or
    nan and the search goes on with that mu).  With no point left: GuessError on
    the first step (i = 0), JError after it."""
    i = len(r) - 2 if i is None else i
    hi = 10.0 * max(mu)
    j = i
    while True:
        den = r[i + 1] - r[j]
        new = (r[i + 1] * mu[j] - r[j] * mu[i + 1]) / den if den != 0 else float("nan")
        if den == 0 or (hi <= new <= 0.0):
            j -= 1
            if j < 0:
                if i == 0:
                    raise GuessError("Bad guess.")
                raise JError("j too small. Secant method failed")
            continue
        return new


def check_to_stop(residual: float, err: float) -> bool:
    """mylib.py:582-584: stop once |observable - target| < err."""
    return abs(residual) < err


def change_guess(r: Sequence[float], guess: float) -> float:
    """change_guess (mylib.py:591-593): both initial residuals positive halves
    the guess.  Otherwise the reference returns None and its caller fails on
    unpacking it (and it also unpacks the halved guess into (guess, delta));
    here delta is kept and the other case raises GuessError."""
    if r[0] > 0 and r[1] > 0:
        return 0.5 * guess
    raise GuessError("Bad guess: the initial residuals do not both lie above the target")


def _value_err(v):
    return (float(v[0]), float(v[1])) if isinstance(v, (tuple, list)) else (float(v), None)


def secant_mu_search(run_density: Callable[[float], float], mu0: float, mu1: float,
                     target: float = 1.0, tol: float = 1e-3, max_iter: int = 15):
    """Secant iteration of mu with density(mu) = target, each evaluation a
    server run.  run_density returns the density or (density, standard
    error); with an error the search stops when the residual is inside it
    (check_to_stop(rho - 1, rho_err), mylib.py:687), else inside `tol`.
    Returns (mu, history of (mu, residual))."""
    mus, res = [mu0, mu1], []
    errs = []
    for m in mus:
        v, e = _value_err(run_density(m))
        res.append(v - target)
        errs.append(tol if e is None else e)
    for _ in range(max_iter):
        if check_to_stop(res[-1], errs[-1]):
            break
        m = new_mu(res, mus)
        mus.append(m)
        v, e = _value_err(run_density(m))
        res.append(v - target)
        errs.append(tol if e is None else e)
    return mus[-1], list(zip(mus, res))


def secant_method(run_many: Callable[[Sequence[float]], Sequence[tuple]], guess: float,
                  delta: float, target: float = 1.0, max_i: int = 15):
    """The search loop of Secant_method_function (mylib.py:599-706).

    run_many(mus) returns one (density, standard error) per mu; the first two
    points mu = round((1 +- delta) guess, 5) are run together, as the reference
    runs them in two threads.  A GuessError on the first step restarts from
    change_guess; a JError later ends the search; the search stops when the
    last density is within its error of the target (check_to_stop) or at
    i == 15 as in the reference.  Returns (mu, history of (mu, density, err),
    converged)."""
    mu_s = [round((1 + delta) * guess, 5), round((1 - delta) * guess, 5)]
    rhos: list = []
    errs: list = []
    hist: list = []
    found = None
    i, restarts = 0, 0
    while True:
        if i == 0:
            (r0, e0), (r1, e1) = run_many(mu_s[:2])
            rhos, errs = [r0 - target, r1 - target], [e0, e1]
            hist += [(mu_s[0], r0, e0), (mu_s[1], r1, e1)]
            if check_to_stop(rhos[0], e0):
                found = mu_s[0]
            if check_to_stop(rhos[1], e1):
                found = mu_s[1]
            try:
                mu_s.append(round(new_mu(rhos, mu_s, 0), 5))
            except GuessError:
                restarts += 1
                if restarts > 8:  # the reference would halve the guess forever
                    raise
                guess = change_guess(rhos, guess)
                mu_s = [round((1 + delta) * guess, 5), round((1 - delta) * guess, 5)]
                continue
            i += 1
        else:
            (r, e), = run_many([mu_s[i]])
            rhos.append(r - target)
            errs.append(e)
            hist.append((mu_s[i], r, e))
            try:
                mu_s.append(new_mu(rhos, mu_s, i - 1))
            except JError:
                break
        if check_to_stop(rhos[-1], errs[-1]):
            found = mu_s[len(rhos) - 1]
        if found is not None or i == max_i:
            break
        i += 1
    mu = found if found is not None else mu_s[len(rhos) - 1]
    return mu, hist, found is not None


def windowed_means(series, cut_fraction: float = 0.0, excluded: Sequence[int] = ()):
    """Running means after a burn-in cut, across runs, from the cumulative means
    a run writes at each snapshot (one column of Skalars.txt per run): the
    arithmetic of compute_rhos (mylib.py:384-404).

    series: (runs, rows) cumulative means; i0 = int(cut_fraction * rows).  For
    i = 1 .. rows - i0 - 1 and each run r not in `excluded` (1-based run
    numbers, as the reference's),
        v_r(i) = (i0 + i + 1) / (i + 1) * (m_r[i0 + i] - i0 / (i0 + i + 1) * m_r[i0]),
    the mean over rows i0 .. i0 + i with row k weighted as k + 1 snapshots and
    the start row as i0 (the reference's weights, kept as written).  Returns
    (evol, evol_err): the mean of v_r(i) over runs and its population standard
    deviation / sqrt(runs - len(excluded)).  A NaN in m_r[i0 + i] raises
    ValueError, as the reference raises NanError.  With runs windowed by
    eval_from (clv_request) the server returns such window means directly.
    """
    import numpy as np

    m = np.asarray(series, dtype=float)
    runs, rows = m.shape
    keep = [r for r in range(runs) if r + 1 not in excluded]
    i0 = int(cut_fraction * rows)
    evol, err = [], []
    for i in range(1, rows - i0):
        last = m[keep, i0 + i]
        if np.isnan(last).any():
            raise ValueError("NaN at line %d" % (i0 + i))
        v = (i0 + (i + 1)) / (i + 1) * (last - i0 / (i0 + (i + 1)) * m[keep, i0])
        evol.append(float(np.mean(v)))
        err.append(float(np.var(v) ** 0.5 / (runs - len(excluded)) ** 0.5))
    return np.array(evol), np.array(err)
