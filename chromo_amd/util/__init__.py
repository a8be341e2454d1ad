"""Host utilities: the ssWLC parameter table and small helpers.

``dss_params`` is the discretised-wormlike-chain parameter table shipped with the
reference (chromo/util/dssWLCparams, loaded at chromo/util/__init__.py:5-6); it is
data, kept byte-identical so that ``SSWLC._find_parameters`` interpolates the same
numbers.  Columns: Delta, eps_b, gamma, eps_par, eps_perp, eta, (two unused).
"""
from pathlib import Path

import numpy as np

_util_folder = Path(__file__).parent.absolute()
dss_params = np.loadtxt(_util_folder / Path("dssWLCparams"))
