"""RV data staging: the reference's ``DataWrapper`` semantics without pandas.

Follows emp/qol_utils.py:16-100 (``DataWrapper.activate``/``mk_RV``) and
emp/qol_utils.py:192-198 (``get_data__``):

* files are read from ``<read_loc>datafiles/<target>/RV/`` in sorted
  filename order (``.DS_Store`` skipped); the 1-based file index is the
  instrument ``Flag``;
* each file's RV mean is subtracted when ``|mean| > 1e-6``;
* extra columns are stellar-activity indices (qol_utils.py:73-90): each is
  mean-subtracted, then mapped linearly onto the file's RV range
  ``(x - min x) / (max x - min x) * (max RV - min RV) + min RV``; after the
  concatenation an epoch holds zeros in the other files' columns
  (``fillna(0)``, qol_utils.py:198).  ``RVData.SA`` is that [N, sum(nsai)]
  matrix, ``RVData.nsai`` the count per file; the Simulation keeps them only
  with ``switch_SA`` (emp/emp.py:2305-2312);
* all epochs are concatenated, sorted by BJD and shifted by
  ``common_t = min(BJD)``.
"""
from __future__ import annotations

import os

import numpy as np


def _scale_activity(col, rv):
    """qol_utils.py:86-90: mean-subtract, then map [min, max] onto the file's
    RV range.  A constant column gives 0/0; the loader's fillna(0)
    (qol_utils.py:198) turns those NaNs into zeros."""
    col = col - col.mean()
    lo, hi = col.min(), col.max()
    with np.errstate(invalid="ignore", divide="ignore"):
        out = (col - lo) / (hi - lo) * (rv.max() - rv.min()) + rv.min()
    return np.where(np.isnan(out), 0.0, out)


class RVData:
    """Time-sorted RV data set: arrays X (BJD - common_t), Y, YERR, FLAG
    (1-based), and the stellar-activity columns SA [N, sum(nsai)] with their
    count per instrument nsai."""

    def __init__(self, X, Y, YERR, FLAG, common_t=0.0, labels=None, SA=None, nsai=None):
        self.X = np.ascontiguousarray(X, dtype=np.float64)
        self.Y = np.ascontiguousarray(Y, dtype=np.float64)
        self.YERR = np.ascontiguousarray(YERR, dtype=np.float64)
        self.FLAG = np.ascontiguousarray(FLAG, dtype=np.int64)
        self.common_t = float(common_t)
        self.nins = int(self.FLAG.max()) if self.FLAG.size else 0
        self.labels = labels or [f"ins{i + 1}" for i in range(self.nins)]
        if not (len(self.X) == len(self.Y) == len(self.YERR) == len(self.FLAG)):
            raise ValueError("X, Y, YERR and FLAG must have the same length")
        if self.nins and set(np.unique(self.FLAG)) != set(range(1, self.nins + 1)):
            raise ValueError("FLAG must use every instrument index 1..nins")
        self.nsai = [0] * self.nins if nsai is None else [int(v) for v in nsai]
        if len(self.nsai) != self.nins:
            raise ValueError("nsai needs one count per instrument")
        S = sum(self.nsai)
        self.SA = (np.zeros((len(self.X), 0)) if SA is None
                   else np.ascontiguousarray(SA, dtype=np.float64).reshape(len(self.X), -1))
        if self.SA.shape[1] != S:
            raise ValueError(f"SA has {self.SA.shape[1]} columns, nsai sums to {S}")

    def without_activity(self):
        """The same epochs without activity columns (switch_SA off)."""
        return RVData(self.X, self.Y, self.YERR, self.FLAG, self.common_t, self.labels)

    @property
    def ndat(self):
        return len(self.X)

    def as_table(self):
        """Dict view with the reference's column names (BJD, RV, eRV, Flag)."""
        return {"BJD": self.X, "RV": self.Y, "eRV": self.YERR, "Flag": self.FLAG}

    @classmethod
    def from_arrays(cls, bjd_list, rv_list, erv_list, labels=None, sa_list=None):
        """Build from per-instrument arrays, applying the loader's rules.
        sa_list: per instrument an [n_k, m_k] array of raw activity indices
        (m_k may be 0), or None."""
        frames = []
        nsai = []
        for k, (b, r, e) in enumerate(zip(bjd_list, rv_list, erv_list)):
            b = np.asarray(b, dtype=np.float64)
            r = np.asarray(r, dtype=np.float64)
            e = np.asarray(e, dtype=np.float64)
            mean = r.mean()
            if abs(mean) > 1e-6:
                r = r - mean
            sa = np.zeros((len(b), 0)) if sa_list is None or sa_list[k] is None else \
                np.asarray(sa_list[k], dtype=np.float64).reshape(len(b), -1)
            sa = np.stack([_scale_activity(sa[:, j], r) for j in range(sa.shape[1])], axis=1) \
                if sa.shape[1] else sa
            nsai.append(sa.shape[1])
            frames.append((b, r, e, np.full(len(b), k + 1, dtype=np.int64), sa))
        bjd = np.concatenate([f[0] for f in frames])
        rv = np.concatenate([f[1] for f in frames])
        erv = np.concatenate([f[2] for f in frames])
        flag = np.concatenate([f[3] for f in frames])
        # columns instrument-major, zeros at other instruments' epochs
        SA = np.zeros((len(bjd), sum(nsai)))
        row, col = 0, 0
        for f, m in zip(frames, nsai):
            SA[row:row + len(f[0]), col:col + m] = f[4]
            row += len(f[0])
            col += m
        order = np.argsort(bjd, kind="quicksort")
        bjd, rv, erv, flag, SA = bjd[order], rv[order], erv[order], flag[order], SA[order]
        common_t = bjd.min()
        return cls(bjd - common_t, rv, erv, flag, common_t, labels, SA=SA, nsai=nsai)


class DataWrapper:
    """Reads ``datafiles/<target_name>/RV/*`` like emp/qol_utils.py:16-100."""

    def __init__(self, target_name, read_loc=""):
        self.target_name = target_name
        self.path = os.path.join(read_loc, "datafiles", target_name, "RV")
        if not os.path.isdir(self.path):
            raise FileNotFoundError(f"no RV folder at {self.path}")
        self.filenames = [fn for fn in sorted(os.listdir(self.path)) if fn != ".DS_Store"]
        if not self.filenames:
            raise FileNotFoundError(f"no RV files in {self.path}")
        bjds, rvs, ervs, sas = [], [], [], []
        self.ndata, self.ncols = [], []
        for fn in self.filenames:
            d = np.loadtxt(os.path.join(self.path, fn))
            if d.ndim == 1:
                d = d[None, :]
            self.ndata.append(d.shape[0])
            self.ncols.append(d.shape[1])
            bjds.append(d[:, 0])
            rvs.append(d[:, 1])
            ervs.append(d[:, 2])
            sas.append(d[:, 3:])
        self.data = RVData.from_arrays(bjds, rvs, ervs, labels=list(self.filenames), sa_list=sas)
        self.nsai = list(self.data.nsai)                     # qol_utils.py:76-79

    def get_data__(self):
        return self.data
