"""Statistical comparators for the tests: the package's own (mcrt.stats) plus
the oracle's raw output layout."""
from __future__ import annotations

import numpy as np

from mcrt.stats import compare, flat  # noqa: F401


def oracle_flat(res):
    return np.concatenate([res["A"], np.asarray(res["Irt"]).transpose(2, 1, 0).ravel(),
                           np.asarray(res["Icb"]).transpose(2, 1, 0).ravel()])
