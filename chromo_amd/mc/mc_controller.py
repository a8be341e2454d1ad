"""Move amplitude control (chromo/mc/mc_controller.py:16-468).

The controllers' arithmetic runs on the device after every controller block of
an ``mc_sim`` step (the same recurrences as ``SimpleControl``); these classes
carry the configuration and receive the state back after each call.
"""
from __future__ import annotations

from typing import Dict, List, Optional, Tuple

from . import moves as mv


class Controller:
    """Controller of one MCAdapter (mc_controller.py:16-75)."""

    def __init__(self, mc_adapter: mv.MCAdapter, bead_amp_bounds: Tuple[int, int],
                 move_amp_bounds: Tuple[float, float]):
        if bead_amp_bounds[0] > bead_amp_bounds[1]:
            raise ValueError("Lower bead amplitude bound must be less than upper bead amplitude bound")
        if move_amp_bounds[0] > move_amp_bounds[1]:
            raise ValueError("Lower move amplitude bound must be less than upper move amplitude bound")
        self.move = mc_adapter
        self.bead_amp_bounds = bead_amp_bounds
        self.move_amp_bounds = move_amp_bounds

    def update_amplitudes(self):
        pass


class NoControl(Controller):
    """Keeps amplitudes fixed (mc_controller.py:78-94)."""

    def update_amplitudes(self):
        return


class SimpleControl(Controller):
    """Factor-0.95 move amplitude control around a 0.5 acceptance setpoint,
    stepping the bead amplitude at the move-amplitude bounds (mc_controller.py:97-213)."""

    def __init__(self, mc_adapter, bead_amp_bounds, move_amp_bounds):
        super().__init__(mc_adapter=mc_adapter, bead_amp_bounds=bead_amp_bounds,
                         move_amp_bounds=move_amp_bounds)
        self.name = "SimpleController"

    def to_file(self, path):
        pass

    def update_amplitudes(self, setpoint_acceptance: Optional[float] = 0.5,
                          move_adjust_factor: Optional[float] = 0.95,
                          num_delta_beads: Optional[int] = 1):
        """Host form of the update the device applies (used outside mc_sim)."""
        acc = self.move.acceptance_tracker.acceptance_rate
        if acc < setpoint_acceptance:
            prop = self.move.amp_move * move_adjust_factor
            if prop > self.move_amp_bounds[0]:
                self.move.amp_move = prop
            else:
                self.move.amp_move = self.move_amp_bounds[1]
                self.move.amp_bead = int(max(self.bead_amp_bounds[0],
                                             self.move.amp_bead - num_delta_beads))
        elif acc > setpoint_acceptance:
            prop = self.move.amp_move / move_adjust_factor
            if prop < self.move_amp_bounds[1]:
                self.move.amp_move = prop
            else:
                self.move.amp_move = self.move_amp_bounds[0]
                self.move.amp_bead = int(min(self.bead_amp_bounds[1],
                                             self.move.amp_bead + num_delta_beads))


def _controllers(log_dir, bead_amp_bounds, move_amp_bounds, controller, kinds, per_cycle):
    out = []
    for kind, npc in zip(kinds, per_cycle):
        ad = mv.MCAdapter(str(log_dir) + '/acceptance_trackers', kind.__name__ + "_snap_", kind,
                          moves_in_average=20, init_amp_bead=bead_amp_bounds[kind.__name__][0],
                          init_amp_move=move_amp_bounds[kind.__name__][0])
        ad.num_per_cycle = npc
        out.append(controller(ad, bead_amp_bounds=bead_amp_bounds[kind.__name__],
                              move_amp_bounds=move_amp_bounds[kind.__name__]))
    return out


def all_moves(log_dir: str, bead_amp_bounds: Dict[str, Tuple[float, float]],
              move_amp_bounds: Dict[str, Tuple[float, float]],
              controller=NoControl) -> List[Controller]:
    """Crank-shaft x30, end-pivot x1, slide x60, tangent rotation x60, binding x10
    per step (mc_controller.py:216-266)."""
    return _controllers(log_dir, bead_amp_bounds, move_amp_bounds, controller,
                        [mv.crank_shaft, mv.end_pivot, mv.slide, mv.tangent_rotation,
                         mv.change_binding_state], [30, 1, 60, 60, 10])


def all_moves_except_binding_state(log_dir: str, bead_amp_bounds, move_amp_bounds,
                                   controller=NoControl) -> List[Controller]:
    """The four geometric moves, 5 each per step (mc_controller.py:269-318)."""
    return _controllers(log_dir, bead_amp_bounds, move_amp_bounds, controller,
                        [mv.crank_shaft, mv.end_pivot, mv.slide, mv.tangent_rotation],
                        [5, 5, 5, 5])


def specific_move(move_fxn, log_dir: str, bead_amp_bounds, move_amp_bounds,
                  controller=NoControl) -> List[Controller]:
    """A single move, once per step (mc_controller.py:321-360)."""
    return _controllers(log_dir, bead_amp_bounds, move_amp_bounds, controller, [move_fxn], [1])
