"""Chain storage: the run's history in one ``.npz`` file.

The reference saves reddemcee's PTHDFBackend (emp/emp.py:722-762): per
temperature ``chain`` [R, W, D], ``log_like`` and ``log_prob`` [R, W],
``beta_history`` [R] and ``accepted`` [W] (R rows = sweeps x nsteps), and
``tsw_history`` / ``smd_history`` [sweeps, T-1] with the ``iteration``
count.  h5py is not part of this stack, so the same arrays
(temperature-major) go into a NumPy archive; :class:`SavedRun` reads it back
with the getters' names and their ``discard`` / ``thin`` / ``flat`` arguments
(emp/emp.py:1401-1409).
"""
from __future__ import annotations

import numpy as np

FORMAT = "astroemperor_amd.ptrun/1"


def save_run(sampler, path):
    """Writes the sampler's recorded history (every sweep) to `path`."""
    chain = sampler.get_chain()                       # [T, S, W, D]
    np.savez(path, format=np.array(FORMAT), iteration=np.array(sampler.iteration, dtype=np.int64),
             chain=chain, log_like=sampler.get_log_like(), log_prob=sampler.get_log_prob(),
             beta_history=np.moveaxis(sampler.get_betas(), 0, 1), accepted=sampler._get(6, np.int64)[..., 0, 0],
             tsw_history=sampler.get_tsw(), smd_history=sampler.get_smd(), betas=sampler.betas)


class SavedRun:
    """A saved run (numpy arrays; nothing is executed from the file)."""

    def __init__(self, path):
        with np.load(path, allow_pickle=False) as f:
            if str(f["format"]) != FORMAT:
                raise ValueError(f"{path} is not a saved run of this format")
            self._a = {k: f[k] for k in f.files}
        self.iteration = int(self._a["iteration"])
        self.betas = self._a["betas"]
        self.ntemps, self.nrows, self.nwalkers, self.ndim = self._a["chain"].shape
        self.nsweeps = len(self._a["tsw_history"])

    def _sel(self, arr, discard, thin, flat):
        arr = arr[:, int(discard)::int(thin)]
        if flat:
            s = arr.shape
            arr = arr.reshape(s[0], s[1] * s[2], *s[3:])
        return arr

    def get_chain(self, discard=0, thin=1, flat=False):
        return self._sel(self._a["chain"], discard, thin, flat)

    def get_log_like(self, discard=0, thin=1, flat=False):
        return self._sel(self._a["log_like"], discard, thin, flat)

    def get_log_prob(self, discard=0, thin=1, flat=False):
        return self._sel(self._a["log_prob"], discard, thin, flat)

    def get_betas(self, discard=0, thin=1):
        return np.moveaxis(self._a["beta_history"], 0, 1)[int(discard)::int(thin)]

    def get_tsw(self, discard=0):
        return self._a["tsw_history"][int(discard):]

    def get_smd(self, discard=0):
        return self._a["smd_history"][int(discard):]

    @property
    def accepted(self):
        return self._a["accepted"]

    @property
    def acceptance_fraction(self):
        return self._a["accepted"] / max(self.iteration, 1)

    @property
    def backend(self):
        """Per-temperature views as reddemcee's backend is read (:class:`RunView`)."""
        return RunView(self)


class TemperatureView:
    """One temperature of a run's history, read as reddemcee's per-temperature
    backend ``sampler.backend[t]`` is when EMPEROR copies it to disk
    (emp/emp.py:751-761): ``iteration``, ``get_chain()`` [rows, W, D],
    ``get_log_like()`` / ``get_log_prob()`` [rows, W], ``get_betas()`` [rows]
    (the ladder of each row's sweep) and ``accepted`` [W].  Rows are stretch
    steps (sweeps x nsteps)."""

    def __init__(self, run, t):
        self._run, self._t = run, int(t)

    @property
    def iteration(self):
        return self._run.iteration

    def get_chain(self, discard=0, thin=1, flat=False):
        out = self._run.get_chain(discard, thin)[self._t]
        return out.reshape(-1, out.shape[-1]) if flat else out

    def get_log_like(self, discard=0, thin=1, flat=False):
        out = self._run.get_log_like(discard, thin)[self._t]
        return out.ravel() if flat else out

    def get_log_prob(self, discard=0, thin=1, flat=False):
        out = self._run.get_log_prob(discard, thin)[self._t]
        return out.ravel() if flat else out

    def get_betas(self, discard=0, thin=1):
        return self._run.get_betas(discard, thin)[:, self._t]

    @property
    def accepted(self):
        return np.asarray(self._run.accepted)[self._t]


class RunView:
    """The ensemble-level backend ``sampler.backend`` as EMPEROR reads it
    (emp/emp.py:734-747): ``iteration``, ``tsw_history_bool`` /
    ``smd_history_bool`` and ``tsw_history`` / ``smd_history`` [sweeps, T-1];
    ``backend[t]`` is a :class:`TemperatureView`."""

    def __init__(self, run):
        self._run = run

    @property
    def iteration(self):
        return self._run.iteration

    @property
    def tsw_history_bool(self):
        return bool(getattr(self._run, "tsw_history_bool", True))

    @property
    def smd_history_bool(self):
        return bool(getattr(self._run, "smd_history_bool", True))

    @property
    def tsw_history(self):
        return self._run.get_tsw()

    @property
    def smd_history(self):
        return self._run.get_smd()

    def __len__(self):
        return self._run.ntemps

    def __getitem__(self, t):
        if not -self._run.ntemps <= int(t) < self._run.ntemps:
            raise IndexError(f"temperature {t} of {self._run.ntemps}")
        return TemperatureView(self._run, int(t) % self._run.ntemps)
