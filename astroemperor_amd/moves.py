"""Moves a :class:`PTSampler` sweep can make.

reddemcee.PTSampler takes emcee's ``moves=[(move, weight), ...]``; EMPEROR's
engine configuration leaves it at None (emp/emp.py:2373), i.e. the stretch
move alone, which stays the default here.  Each sweep makes one kind of move at
every temperature, chosen with probability weight / sum(weights).
"""
import numpy as np

from . import _abi


class StretchMove:
    """Goodman & Weare (2010) stretch move: Y = X_j + z (X_k - X_j) with
    z = ((a - 1) u + 1)^2 / a (emcee's StretchMove, a = 2 by default)."""

    kind = _abi.EMP_MOVE_STRETCH

    def __init__(self, a=2.0):
        self.a = float(a)

    def scale(self, ndim):
        return self.a

    def __repr__(self):
        return f"StretchMove(a={self.a})"


class GradientMove:
    """Ensemble-preconditioned Metropolis-adjusted Langevin move.

    Walker x of one half at inverse temperature beta proposes
    ``y = x + L ((step^2 / 2) beta L^T grad lnL(x) + step xi)``, xi ~ N(0, I),
    where L L^T is the covariance of the other half's walkers at the same
    temperature (Leimkuhler, Matthews & Weare 2018), and accepts with the exact
    Metropolis-Hastings ratio of beta lnL + ln prior (Roberts & Tweedie 1996).
    Gradients come from the adjoint kernel.  The default step, 1.65
    ndim^(-1/6), is Roberts & Rosenthal's (1998) optimal MALA scale for a
    target whose covariance the preconditioner matches.  ``step`` may also
    be one value per temperature (hottest last, as the ladder), e.g. the
    steps :meth:`PTSampler.adapt_gradient_steps` found.
    """

    kind = _abi.EMP_MOVE_GRADIENT

    def __init__(self, step=None):
        if step is None or np.ndim(step) == 0:
            self.step = None if step is None else float(step)
            self.steps = None
        else:
            self.steps = np.ascontiguousarray(step, dtype=np.float64)
            if self.steps.ndim != 1 or not np.all((self.steps > 0) & np.isfinite(self.steps)):
                raise ValueError("per-temperature steps must be a 1-D array of finite values > 0")
            self.step = float(self.steps[0])

    def scale(self, ndim):
        return self.step if self.step is not None else 1.65 * float(ndim) ** (-1.0 / 6.0)

    def __repr__(self):
        if self.steps is not None:
            return f"GradientMove(steps={np.array2string(self.steps, precision=3)})"
        return f"GradientMove(step={self.step})"


_MOVES = (StretchMove, GradientMove)


def normalise_moves(moves, a=2.0):
    """emcee's forms: None, one move, a list of moves (equal weights) or a
    list of (move, weight) pairs.  Returns [(move, weight)]."""
    if moves is None:
        return [(StretchMove(a), 1.0)]
    if isinstance(moves, _MOVES):
        return [(moves, 1.0)]
    out = []
    for item in moves:
        if isinstance(item, _MOVES):
            out.append((item, 1.0))
        else:
            m, w = item
            if not isinstance(m, _MOVES):
                raise TypeError(f"unsupported move {m!r}: StretchMove or GradientMove")
            out.append((m, float(w)))
    kinds = [m.kind for m, _ in out]
    if not out or len(set(kinds)) != len(kinds):
        raise ValueError("one or two moves, each kind at most once")
    if any(not (w >= 0) for _, w in out) or sum(w for _, w in out) <= 0:
        raise ValueError("move weights must be >= 0 with a positive sum")
    return out


def move_descs(moves, ndim):
    arr = (_abi.EmpMoveDesc * len(moves))()
    for i, (m, w) in enumerate(moves):
        arr[i].kind, arr[i].weight, arr[i].scale = m.kind, w, m.scale(ndim)
    return arr
