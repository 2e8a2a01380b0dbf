"""Serving: continuous batching of Langevin runs on one GPU, and the secant
mu search of the reference's orchestration as a client of it.

`Server` wraps clv_server_* (include/clv.h): runs are submitted at any time,
admitted into device-resident slot pools by HBM budget while others are in
flight, and return their observables and time-to-result.  `secant_mu_search`
restates the secant update of mylib.py:561-584 (new_mu / check_to_stop) with
follow-up runs issued to the server toward a target density; `windowed_means`
restates the burn-in-cut running means of mylib.py:384-404.
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


def new_mu(r: Sequence[float], mu: Sequence[float]) -> float:
    """Secant step mu_{n+1} = (r_{n+1} mu_n - r_n mu_{n+1}) / (r_{n+1} - r_n) on
    the residuals r = observable - target (mylib.py:561-580)."""
    r0, r1 = r[-2], r[-1]
    m0, m1 = mu[-2], mu[-1]
    if r1 == r0:
        return m1
    return (r1 * m0 - r0 * m1) / (r1 - r0)


def check_to_stop(residual: float, err: float) -> bool:
    """mylib.py:582-584: stop once |observable - target| < err."""
    return abs(residual) < err


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


def secant_mu_search(run_density: Callable[[float], float], mu0: float, mu1: float,
                     target: float = 1.0, tol: float = 1e-3, max_iter: int = 15):
    """Secant search of mu with density(mu) = target (the orchestration of
    Secant_method_function, mylib.py:599-706, with each evaluation a server
    run).  Returns (mu, history)."""
    mus = [mu0, mu1]
    res = [run_density(mu0) - target, run_density(mu1) - target]
    for _ in range(max_iter):
        if check_to_stop(res[-1], tol):
            break
        m = new_mu(res, mus)
        mus.append(m)
        res.append(run_density(m) - target)
    return mus[-1], list(zip(mus, res))
