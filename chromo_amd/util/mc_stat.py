"""Acceptance tracking (chromo/util/mc_stat.py:36-208).

Only the exponentially weighted acceptance rate feeds the move controllers; it
is advanced on the device during a run and copied back into these objects.
Move logging (``log_move``/``save_move_log``) keeps the reference's CSV layout.
"""
from __future__ import annotations

import csv
from typing import List


class AcceptanceTracker:
    """EWMA acceptance rate with decay alpha = 2 / (moves_in_average + 1)."""

    def __init__(self, log_dir: str, log_file_prefix: str, moves_in_average: float):
        self.log_dir = log_dir
        self.log_file_prefix = log_file_prefix
        self.alpha = 2 / (moves_in_average + 1)
        self.acceptance_rate = 0
        self.amp_bead_limit_log: List[float] = []
        self.amp_move_limit_log: List[float] = []
        self.amp_bead_realized_log: List[float] = []
        self.amp_move_realized_log: List[float] = []
        self.dE_log: List[float] = []
        self.move_accepted: List[float] = []
        self.acceptance_log: List[float] = []

    def update_acceptance_rate(self, accept: float, log_update: int = 0):
        """rate <- alpha * accept + (1 - alpha) * rate (mc_stat.py:190-208)."""
        if log_update == 1:
            self.move_accepted.append(accept)
        self.acceptance_rate = (self.alpha * accept) + (1 - self.alpha) * self.acceptance_rate

    def create_log_file(self, ind: int):
        path = f"{self.log_dir}/{self.log_file_prefix}{ind}.csv"
        with open(path, "w", newline="") as f:
            csv.writer(f).writerow([
                "snapshot", "iteration", "bead_amp_limit", "move_amp_limit",
                "bead_amp_realized", "move_amp_realized", "dE", "accepted", "acceptance_rate"])

    def log_move(self, amp_move_limit, amp_bead_limit, amp_move, amp_bead, dE):
        self.amp_move_limit_log.append(amp_move_limit)
        self.amp_bead_limit_log.append(amp_bead_limit)
        self.amp_move_realized_log.append(amp_move)
        self.amp_bead_realized_log.append(amp_bead)
        self.dE_log.append(dE)
        self.acceptance_log.append(self.acceptance_rate)

    def save_move_log(self, snapshot: int):
        n = len(self.amp_move_realized_log)
        if n != len(self.amp_bead_realized_log) or n != len(self.acceptance_log):
            raise ValueError("move log columns have different lengths")
        path = f"{self.log_dir}/{self.log_file_prefix}{snapshot}.csv"
        with open(path, "a", newline="") as f:
            w = csv.writer(f)
            for i in range(n):
                w.writerow([snapshot, i + 1, self.amp_bead_limit_log[i], self.amp_move_limit_log[i],
                            self.amp_bead_realized_log[i], self.amp_move_realized_log[i],
                            self.dE_log[i], self.move_accepted[i] if i < len(self.move_accepted)
                            else "", self.acceptance_log[i]])
        for lst in (self.amp_move_limit_log, self.amp_bead_limit_log, self.amp_move_realized_log,
                    self.amp_bead_realized_log, self.dE_log, self.move_accepted,
                    self.acceptance_log):
            lst.clear()
