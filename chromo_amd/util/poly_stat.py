"""Output-folder helpers of a simulation (chromo/util/poly_stat.py:184-262).

``polymer_in_field`` writes each run into its own ``sim_<i>`` folder of the
output directory, with one CSV snapshot ``<polymer name>-<save point>.csv`` per
polymer and save point.  These helpers find the latest run, its polymers and
their latest snapshots, which ``continue_polymer_in_field_simulation`` resumes
from.
"""
from __future__ import annotations

import os
import re
from typing import List

_SNAPSHOT = re.compile(r"^(?P<name>.+)-(?P<k>\d+)\.csv$")


def get_latest_simulation(directory: str = ".") -> str:
    """The ``sim_<i>`` subfolder of ``directory`` with the largest i."""
    sims = [d for d in os.listdir(directory)
            if os.path.isdir(os.path.join(directory, d)) and re.fullmatch(r"sim_\d+", d)]
    if not sims:
        raise FileNotFoundError(f"no sim_<i> folder in {directory}")
    return max(sims, key=lambda d: int(d.split("_")[-1]))


def find_polymers_in_output_dir(directory: str = ".") -> List[str]:
    """Names of the polymers with snapshots in ``directory``."""
    names = set()
    for f in os.listdir(directory):
        m = _SNAPSHOT.match(f)
        if m and os.path.isfile(os.path.join(directory, f)):
            names.add(m.group("name"))
    return sorted(names)


def get_latest_configuration(polymer_prefix: str = "Chr-1", directory: str = ".") -> str:
    """Path of the snapshot of ``polymer_prefix`` with the largest save point."""
    best = -1
    for f in os.listdir(directory):
        m = _SNAPSHOT.match(f)
        if m and m.group("name") == polymer_prefix:
            best = max(best, int(m.group("k")))
    if best < 0:
        raise FileNotFoundError(f"no snapshot of {polymer_prefix} in {directory}")
    return f"{directory}/{polymer_prefix}-{best}.csv"
