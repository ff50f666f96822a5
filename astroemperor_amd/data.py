"""RV data staging: the reference's ``DataWrapper`` semantics without pandas.

Follows emp/qol_utils.py:16-100 (``DataWrapper.activate``/``mk_RV``) and
emp/qol_utils.py:192-198 (``get_data__``):

* files are read from ``<read_loc>datafiles/<target>/RV/`` in sorted
  filename order (``.DS_Store`` skipped); the 1-based file index is the
  instrument ``Flag``;
* each file's RV mean is subtracted when ``|mean| > 1e-6``;
* extra columns (stellar-activity indices) are read but not used by the RV
  hot path (the reference drops them unless ``switch_SA``, emp/emp.py:2309);
* all epochs are concatenated, sorted by BJD and shifted by
  ``common_t = min(BJD)``.
"""
from __future__ import annotations

import os

import numpy as np


class RVData:
    """Time-sorted RV data set: arrays X (BJD - common_t), Y, YERR, FLAG (1-based)."""

    def __init__(self, X, Y, YERR, FLAG, common_t=0.0, labels=None):
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

    @property
    def ndat(self):
        return len(self.X)

    def as_table(self):
        """Dict view with the reference's column names (BJD, RV, eRV, Flag)."""
        return {"BJD": self.X, "RV": self.Y, "eRV": self.YERR, "Flag": self.FLAG}

    @classmethod
    def from_arrays(cls, bjd_list, rv_list, erv_list, labels=None):
        """Build from per-instrument arrays, applying the loader's rules."""
        frames = []
        for k, (b, r, e) in enumerate(zip(bjd_list, rv_list, erv_list)):
            b = np.asarray(b, dtype=np.float64)
            r = np.asarray(r, dtype=np.float64)
            e = np.asarray(e, dtype=np.float64)
            mean = r.mean()
            if abs(mean) > 1e-6:
                r = r - mean
            frames.append((b, r, e, np.full(len(b), k + 1, dtype=np.int64)))
        bjd = np.concatenate([f[0] for f in frames])
        rv = np.concatenate([f[1] for f in frames])
        erv = np.concatenate([f[2] for f in frames])
        flag = np.concatenate([f[3] for f in frames])
        order = np.argsort(bjd, kind="quicksort")
        bjd, rv, erv, flag = bjd[order], rv[order], erv[order], flag[order]
        common_t = bjd.min()
        return cls(bjd - common_t, rv, erv, flag, common_t, labels)


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
        bjds, rvs, ervs = [], [], []
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
        self.data = RVData.from_arrays(bjds, rvs, ervs, labels=list(self.filenames))

    def get_data__(self):
        return self.data
