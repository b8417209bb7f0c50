"""Fit records and summaries (cnnclustering report.py:13-191)."""

from __future__ import annotations

import time
from collections.abc import MutableSequence


class Record:
    """Cluster result record (report.py:91-191): the fields printed after a fit."""

    __slots__ = ["n_points", "radius_cutoff", "cnn_cutoff", "member_cutoff", "max_clusters",
                 "n_clusters", "ratio_largest", "ratio_noise", "execution_time"]

    def __init__(self, n_points=None, radius_cutoff=None, cnn_cutoff=None, member_cutoff=None,
                 max_clusters=None, n_clusters=None, ratio_largest=None, ratio_noise=None,
                 execution_time=None):
        self.n_points = n_points
        self.radius_cutoff = radius_cutoff
        self.cnn_cutoff = cnn_cutoff
        self.member_cutoff = member_cutoff
        self.max_clusters = max_clusters
        self.n_clusters = n_clusters
        self.ratio_largest = ratio_largest
        self.ratio_noise = ratio_noise
        self.execution_time = execution_time

    def to_dict(self):
        return {k: getattr(self, k) for k in self.__slots__}

    def __repr__(self):
        return f"{type(self).__name__}({', '.join(f'{k}={getattr(self, k)!r}' for k in self.__slots__)})"

    def __str__(self):
        # Same table layout as the reference's Record.__str__ (report.py:142-178):
        # 10-wide left-aligned columns, floats with 3 decimals, hh:mm:s.sss time.
        cells = []
        for attr in self.__slots__[:-1]:
            v = getattr(self, attr)
            if v is None:
                cells.append(f"{v!r:<10}")
            elif isinstance(v, float):
                cells.append(f"{v:<10.3f}")
            else:
                cells.append(f"{v:<10}")
        t = self.execution_time
        if t is None:
            t_str = "None"
        else:
            hours, rest = divmod(t, 3600)
            minutes, seconds = divmod(rest, 60)
            t_str = f"{int(hours):0>2}:{int(minutes):0>2}:{seconds:.3f}"
        header = ("#points   r         c         min       max       "
                  "#clusters %largest  %noise    time     ")
        line = "-" * 95
        return f"{line}\n{header}\n{''.join(cells)}{t_str}\n{line}\n"


class Summary(MutableSequence):
    """List of Records (report.py:13-88)."""

    def __init__(self, iterable=None, record_type=Record):
        self._list = [] if iterable is None else list(iterable)
        self.record_type = record_type

    def __getitem__(self, key):
        return self._list[key]

    def __setitem__(self, key, item):
        self._list[key] = item

    def __delitem__(self, key):
        del self._list[key]

    def __len__(self):
        return len(self._list)

    def insert(self, index, item):
        self._list.insert(index, item)

    def __str__(self):
        return "\n".join(str(r) for r in self._list)

    def to_DataFrame(self):
        """Records as a typed pandas.DataFrame (Summary.to_DataFrame, report.py:55-88):
        nullable integer columns for counts and cutoffs, float64 for the rest."""
        import numpy as np
        import pandas as pd

        dtypes = [pd.Int64Dtype(), np.float64, pd.Int64Dtype(), pd.Int64Dtype(), pd.Int64Dtype(),
                  pd.Int64Dtype(), np.float64, np.float64, np.float64]
        fields = Record.__slots__
        content = [[getattr(r, f) for r in self._list] for f in fields]
        return make_typed_DataFrame(fields, dtypes, content)


def make_typed_DataFrame(columns, dtypes, content=None):
    """pandas.DataFrame with typed columns (report.py:194-210)."""
    import pandas as pd

    assert len(columns) == len(dtypes)
    if content is None:
        content = [[] for _ in columns]
    return pd.DataFrame({k: pd.array(c, dtype=v) for k, v, c in zip(columns, dtypes, content)})


def timed(function):
    """Wall-clock decorator returning (result, seconds) as cluster.pyx:1191-1201."""

    def wrapper(*args, **kwargs):
        go = time.time()
        wrapped = function(*args, **kwargs)
        stop = time.time()
        return wrapped, stop - go

    return wrapper


