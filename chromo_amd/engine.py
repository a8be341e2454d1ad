"""Device state for one or many replicas and the calls into libchromo_hip.so.

A *replica* is one chain in one density field with its own move controllers and
random streams.  ``DeviceReplicas`` keeps every replica's arrays resident in HBM
(structure-of-arrays chains, voxel densities, RNG blocks, controllers, counters,
move arenas) inside a few flat torch buffers, and exposes the C ABI calls
(include/chromo_hip.h) that advance them.  Torch is used only to own device
memory and to order work on the current HIP stream; all arithmetic happens in
the HIP kernels.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib as L


# ---------------------------------------------------------------------------
# Host-side specifications
# ---------------------------------------------------------------------------

@dataclass
class ChainSpec:
    """One chain: coordinates, binder states, marks and per-bond SSWLC parameters."""
    r: np.ndarray                 # (N, 3) float64
    t3: np.ndarray                # (N, 3)
    t2: np.ndarray                # (N, 3)
    states: np.ndarray            # (N, B) int
    marks: np.ndarray             # (N, B) int
    bonds: np.ndarray             # (5, N-1): eps_bend, eps_par, eps_perp, gamma, eta;
                                  # (7, N-1) for a twisting chain: + eps_twist, natural twist
    sites: Sequence[int]          # sites_per_bead per binder
    mu: Sequence[float]           # chemical_potential per binder
    F: np.ndarray                 # (B, n_mark_levels, sites+1) -log single-site Z
    max_binders: int = -1
    mu_adjust: float = 1.0
    twist: bool = False           # SSTWLC (chromo/polymers.pyx:1889-2320)
    bead_length: Optional[np.ndarray] = None     # (N-1,) bond lengths (host only)

    @property
    def bond_rows(self) -> int:
        return 7 if self.twist else 5

    @property
    def N(self) -> int:
        return int(self.r.shape[0])

    @property
    def B(self) -> int:
        return int(self.states.shape[1])


@dataclass
class FieldSpec:
    """UniformDensityField parameters (chromo/fields.pyx:453-712)."""
    field_on: bool = True
    nx: int = 1
    ny: int = 1
    nz: int = 1
    x_width: float = 1.0
    y_width: float = 1.0
    z_width: float = 1.0
    vf_limit: float = 0.5         # stored by the reference as a C float
    chi: float = 1.0
    bead_vol: float = 1.0
    pref: Sequence[float] = ()
    intra: Sequence[float] = ()
    cross: Optional[np.ndarray] = None   # (B, B)
    confine_type: int = 0
    confine_length: float = 0.0
    access_vol: Optional[np.ndarray] = None   # (n_bins,) accessible volumes, None: vol_bin
    fast_field: bool = False                  # sub-bin lookup weights in the move dE
    n_points: int = 1000                      # sub-bins per voxel side (made even)

    @property
    def n_bins(self) -> int:
        return int(self.nx) * int(self.ny) * int(self.nz)


@dataclass
class CtrlSpec:
    """One adaptable move with its controller state (moves.pyx / mc_controller.py)."""
    move: int
    num_per_cycle: int
    amp_move: float
    amp_bead: int
    bead_lo: float
    bead_hi: float
    move_lo: float
    move_hi: float
    on: int = 1
    controlled: int = 1
    rate: float = 0.0
    alpha: float = 2 / (20 + 1)
    num_attempt: int = 0
    num_success: int = 0

    def to_struct(self) -> L.ChromoCtrl:
        return L.ChromoCtrl(self.move, self.on, self.controlled, self.num_per_cycle,
                            float(self.amp_move), int(self.amp_bead), float(self.rate),
                            float(self.alpha), float(self.bead_lo), float(self.bead_hi),
                            float(self.move_lo), float(self.move_hi), int(self.num_attempt),
                            int(self.num_success))

    @classmethod
    def from_struct(cls, c: L.ChromoCtrl) -> "CtrlSpec":
        return cls(move=c.move, num_per_cycle=c.num_per_cycle, amp_move=c.amp_move,
                   amp_bead=c.amp_bead, bead_lo=c.bead_lo, bead_hi=c.bead_hi,
                   move_lo=c.move_lo, move_hi=c.move_hi, on=c.on, controlled=c.controlled,
                   rate=c.rate, alpha=c.alpha, num_attempt=c.num_attempt,
                   num_success=c.num_success)


def field_grid(fs: FieldSpec):
    """Grid quantities exactly as UniformDensityField.init_grid (fields.pyx:546-565)."""
    nx, ny, nz = int(fs.nx), int(fs.ny), int(fs.nz)
    dx = fs.x_width / nx
    dy = fs.y_width / ny
    dz = fs.z_width / nz
    n_bins = nx * ny * nz
    vol_bin = fs.x_width * fs.y_width * fs.z_width / n_bins
    width = (fs.x_width, fs.y_width, fs.z_width)
    half_width = (0.5 * fs.x_width, 0.5 * fs.y_width, 0.5 * fs.z_width)
    half_step = (0.5 * dx, 0.5 * dy, 0.5 * dz)
    return (dx, dy, dz), n_bins, vol_bin, width, half_width, half_step


def default_caps(ctrls: Sequence[CtrlSpec], N: int):
    """(bead_cap, seg_cap): the largest n_inds of any move and of any field move.

    A move selects at most amp_bead (+1 for a crank's clamped end) beads, and the
    controller keeps amp_bead within its upper bound.
    """
    seg = bead = 1
    for c in ctrls:
        if c.on != 1:
            continue
        hi = int(min(max(c.bead_hi, c.amp_bead) + 1, N))
        bead = max(bead, hi)
        if c.move != L.TANGENT_ROTATION:
            seg = max(seg, hi)
    return bead, seg


def _current_stream() -> int:
    return torch.cuda.current_stream().cuda_stream


# ---------------------------------------------------------------------------
# Device replicas
# ---------------------------------------------------------------------------

class DeviceReplicas:
    """Chains, fields, controllers and RNG streams of ``n`` replicas in HBM."""

    CTRL_SLOTS = 8
    SEG_MAX = 32       # beads per move of the one-wavefront evaluator (chromo_seg.hpp)
    WAVE_CHOICES = (1, 2)
    waves = 1              # wavefronts per replica used by run(); see chromo_amd.autotune
    sweeps_per_launch = 1  # sweeps per captured graph launch chosen by chromo_amd.autotune

    def __init__(self, chains: Sequence[ChainSpec], fields: Sequence[FieldSpec],
                 ctrls: Sequence[Sequence[CtrlSpec]], *, rng_mode: int = L.RNG_GLIBC,
                 seeds: Optional[Sequence[int]] = None, streams: Optional[Sequence[int]] = None,
                 caps: Optional[Sequence[tuple]] = None, log_cap: int = 0,
                 device: Optional[torch.device] = None):
        if not torch.cuda.is_available():
            raise RuntimeError("chromo_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.lib = L.lib()
        n = len(chains)
        if n < 1 or len(fields) != n or len(ctrls) != n:
            raise ValueError("chains, fields and ctrls must have the same non-zero length")
        self.n = n
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.chains = list(chains)
        self.fields = list(fields)
        self.rng_mode = int(rng_mode)
        dev = self.device
        Ns = [c.N for c in chains]
        Bs = [c.B for c in chains]
        for c, f in zip(chains, fields):
            if not 1 <= c.B <= L.MAX_BINDERS:
                raise NotImplementedError(f"binder count {c.B} outside 1..{L.MAX_BINDERS}")
            if c.N < 2:
                raise ValueError("a chain needs at least 2 beads")
            if np.asarray(c.bonds).shape != (c.bond_rows, c.N - 1):
                raise ValueError(f"bonds must have shape ({c.bond_rows}, N-1) for this chain")
            if np.any(np.asarray(c.sites) > L.MAX_SITES):
                raise NotImplementedError(f"sites_per_bead above {L.MAX_SITES}")
            if c.marks.size and (c.marks.min() < 0 or c.marks.max() >= L.MAX_MARK):
                raise NotImplementedError(f"chemical_mods values outside 0..{L.MAX_MARK - 1}")
            if f.field_on and f.n_bins >= 2 ** 31:
                raise NotImplementedError("more than 2^31 voxels")
            if f.field_on and max(int(f.nx), int(f.ny), int(f.nz)) > 32767:
                raise NotImplementedError("more than 32767 voxels along one axis")
        self.Ns, self.Bs = Ns, Bs
        # bit B: B binder types present; bit 8 + B: twisting chains among them
        self.binder_mask = (sum(1 << b for b in set(Bs)) |
                            sum(1 << (8 + c.B) for c in chains if c.twist))
        nctrl = [len(c) for c in ctrls]
        if max(nctrl) > self.CTRL_SLOTS:
            raise NotImplementedError(f"at most {self.CTRL_SLOTS} controllers per replica")
        self.nctrl = list(nctrl)
        if caps is None:
            caps = [default_caps(c, N) for c, N in zip(ctrls, Ns)]
        # every replica gets a global arena of at least min(N, 32) beads: moves of up
        # to 32 beads fall back to it when they touch more voxels than the
        # registers of the one-wavefront evaluator hold
        self.bead_cap = [max(int(c[0]), min(N, self.SEG_MAX)) for c, N in zip(caps, Ns)]
        self.seg_cap = [min(max(int(c[1]), min(N, self.SEG_MAX)), bc)
                        for c, N, bc in zip(caps, Ns, self.bead_cap)]

        # offsets into flat buffers
        def offsets(sizes):
            o = np.zeros(len(sizes) + 1, dtype=np.int64)
            o[1:] = np.cumsum(sizes)
            return o

        self.o_pos = offsets([9 * N for N in Ns])
        self.o_st = offsets([B * N for N, B in zip(Ns, Bs)])
        self.o_bond = offsets([c.bond_rows * (N - 1) for c, N in zip(chains, Ns)])
        self.twist = any(c.twist for c in chains)
        nbins = [f.n_bins if f.field_on else 1 for f in fields]
        self.nbins = nbins
        self.o_den = offsets([nb * (B + 1) for nb, B in zip(nbins, Bs)])
        self.o_ctrl = offsets([self.CTRL_SLOTS] * n)
        sbytes = [int(self.lib.chromo_scratch_bytes(bc, sc, B))
                  for bc, sc, B in zip(self.bead_cap, self.seg_cap, Bs)]
        if any(b < 0 for b in sbytes):
            raise ValueError("invalid arena capacities")
        sbytes = [(b + 255) // 256 * 256 for b in sbytes]
        self.o_scr = offsets(sbytes)
        self.log_cap = int(log_cap)

        # host staging
        pos = np.empty(int(self.o_pos[-1]), dtype=np.float64)
        st = np.empty(max(int(self.o_st[-1]), 1), dtype=np.int8)
        mk = np.empty(max(int(self.o_st[-1]), 1), dtype=np.int8)
        bo = np.empty(max(int(self.o_bond[-1]), 1), dtype=np.float64)
        for i, c in enumerate(chains):
            N, B = Ns[i], Bs[i]
            p = pos[self.o_pos[i]:self.o_pos[i + 1]].reshape(9, N)
            p[0:3] = np.asarray(c.r, dtype=np.float64).T
            p[3:6] = np.asarray(c.t3, dtype=np.float64).T
            p[6:9] = np.asarray(c.t2, dtype=np.float64).T
            st[self.o_st[i]:self.o_st[i + 1]] = np.asarray(c.states).T.reshape(-1).astype(np.int8)
            mk[self.o_st[i]:self.o_st[i + 1]] = np.asarray(c.marks).T.reshape(-1).astype(np.int8)
            bo[self.o_bond[i]:self.o_bond[i + 1]] = np.asarray(c.bonds, dtype=np.float64).reshape(-1)
        self.pos = torch.from_numpy(pos).to(dev)
        self.states = torch.from_numpy(st).to(dev)
        self.marks = torch.from_numpy(mk).to(dev)
        self.bonds = torch.from_numpy(bo).to(dev)
        self.density = torch.zeros(int(self.o_den[-1]), dtype=torch.float64, device=dev)
        # accessible volumes of the fields that have them (spherical confinement
        # without assume_fully_accessible), one flat buffer
        av_list = []
        for f in fields:
            a = getattr(f, "access_vol", None) if f.field_on else None
            if a is not None:
                a = np.asarray(a, dtype=np.float64).reshape(-1)
                if a.size != f.n_bins:
                    raise ValueError("access_vol must hold one volume per voxel")
            av_list.append(a)
        self.o_av = offsets([0 if a is None else a.size for a in av_list])
        self.access_vol = (torch.from_numpy(np.concatenate([a for a in av_list if a is not None])).to(dev)
                           if any(a is not None for a in av_list) else None)
        self.rng = torch.zeros(n * L.RNG_WORDS, dtype=torch.int32, device=dev)
        cs = ctypes.sizeof(L.ChromoCtrl)
        ctrl_host = (L.ChromoCtrl * int(self.o_ctrl[-1]))()
        for i, cl in enumerate(ctrls):
            for j, c in enumerate(cl):
                ctrl_host[self.CTRL_SLOTS * i + j] = c.to_struct()
        self.ctrl = torch.frombuffer(bytearray(ctrl_host), dtype=torch.uint8).to(dev)
        self.counters = torch.zeros(n * L.CNT_WORDS, dtype=torch.int64, device=dev)
        self.scratch = torch.zeros(max(int(self.o_scr[-1]), 1), dtype=torch.uint8, device=dev)
        ls = ctypes.sizeof(L.ChromoMoveLog)
        self.log = (torch.zeros(max(n * self.log_cap * ls, 1), dtype=torch.uint8, device=dev)
                    if self.log_cap > 0 else None)

        # replica table
        tab = (L.ChromoReplica * n)()
        base = lambda t: t.data_ptr()
        for i in range(n):
            c, f, R = chains[i], fields[i], tab[i]
            N, B = Ns[i], Bs[i]
            R.pos = base(self.pos) + 8 * int(self.o_pos[i])
            R.states = base(self.states) + int(self.o_st[i])
            R.marks = base(self.marks) + int(self.o_st[i])
            R.bonds = base(self.bonds) + 8 * int(self.o_bond[i])
            R.density = base(self.density) + 8 * int(self.o_den[i])
            R.rng = base(self.rng) + 4 * L.RNG_WORDS * i
            R.ctrl = base(self.ctrl) + cs * int(self.o_ctrl[i])
            R.log = (base(self.log) + ls * self.log_cap * i) if self.log is not None else None
            R.counters = base(self.counters) + 8 * L.CNT_WORDS * i
            R.scratch = (base(self.scratch) + int(self.o_scr[i])) if sbytes[i] else None
            R.N = N
            R.log_cap = self.log_cap
            R.seg_cap = self.seg_cap[i] if sbytes[i] else 0
            R.bead_cap = self.bead_cap[i] if sbytes[i] else 0
            R.B = B
            R.n_ctrl = nctrl[i]
            R.rng_mode = self.rng_mode
            R.twist = 1 if c.twist else 0
            R.field_on = 1 if f.field_on else 0
            if f.field_on:
                (dx, dy, dz), nb, vol_bin, width, hw, hs = field_grid(f)
                R.nx, R.ny, R.nz = int(f.nx), int(f.ny), int(f.nz)
                for j in range(3):
                    R.width[j] = width[j]
                    R.dxyz[j] = (dx, dy, dz)[j]
                    R.half_width[j] = hw[j]
                    R.half_step[j] = hs[j]
                R.vol_bin = vol_bin
                R.chi = float(f.chi)
                R.bead_vol = float(f.bead_vol)
                R.vf_limit = float(np.float32(f.vf_limit))
                cross = np.zeros((B, B)) if f.cross is None else np.asarray(f.cross, dtype=float)
                for a in range(B):
                    R.pref[a] = float(f.pref[a])
                    R.intra[a] = float(f.intra[a])
                    for b2 in range(B):
                        R.cross[a * B + b2] = float(cross[a, b2])
                R.confine_type = int(f.confine_type)
                R.confine_length = float(f.confine_length)
                if av_list[i] is not None:
                    R.access_vol = base(self.access_vol) + 8 * int(self.o_av[i])
                if f.fast_field:
                    npts = int(f.n_points) + int(f.n_points) % 2      # init_fast_field (fields.pyx:610)
                    if npts < 2:
                        raise ValueError("fast_field needs n_points >= 1")
                    R.fast_field = 1
                    R.half_points = float(npts // 2)
                    for j, (nn, d) in enumerate(zip((f.nx, f.ny, f.nz), (dx, dy, dz))):
                        R.n_sub[j] = npts * int(nn)
                        R.sub_width[j] = d / npts
            else:
                R.nx = R.ny = R.nz = 1
            for a in range(B):
                R.mu[a] = float(c.mu[a])
                R.sites[a] = int(c.sites[a])
            R.mu_adjust = float(c.mu_adjust)
            R.max_binders = int(c.max_binders)
            F = np.asarray(c.F, dtype=np.float64)
            R.n_mark_levels = int(F.shape[1])
            for a in range(B):
                for m in range(min(F.shape[1], L.MAX_MARK)):
                    for s in range(min(F.shape[2], L.MAX_SITES + 1)):
                        R.F[(a * L.MAX_MARK + m) * (L.MAX_SITES + 1) + s] = float(F[a, m, s])
        self._tab_host = tab
        self.table = torch.frombuffer(bytearray(tab), dtype=torch.uint8).to(dev)
        self.tab_ptr = self.table.data_ptr()

        stream = _current_stream()
        L.check(self.lib.chromo_scratch_init(self.tab_ptr, n, stream), "chromo_scratch_init")
        self.seed(seeds if seeds is not None else [1] * n, streams)
        if any(f.field_on for f in fields):
            L.check(self.lib.chromo_density_init(self.tab_ptr, n, stream), "chromo_density_init")

    # -- RNG -----------------------------------------------------------------
    def seed(self, seeds: Sequence[int], streams: Optional[Sequence[int]] = None) -> None:
        """Seed every replica: srand(seed) (glibc mode) or Philox key (seed, stream)."""
        s = torch.tensor(np.asarray(seeds, dtype=np.uint32).view(np.int32), device=self.device)
        st = (torch.tensor(np.asarray(streams, dtype=np.uint32).view(np.int32), device=self.device)
              if streams is not None else None)
        L.check(self.lib.chromo_rng_seed(self.tab_ptr, self.n, s.data_ptr(),
                                         st.data_ptr() if st is not None else None,
                                         self.rng_mode, _current_stream()), "chromo_rng_seed")
        torch.cuda.synchronize(self.device)

    def np_seed(self, seeds: Sequence[int]) -> None:
        s = torch.tensor(np.asarray(seeds, dtype=np.uint32).view(np.int32), device=self.device)
        L.check(self.lib.chromo_np_seed(self.tab_ptr, self.n, s.data_ptr(), _current_stream()),
                "chromo_np_seed")
        torch.cuda.synchronize(self.device)

    def rng_words(self, i: int) -> np.ndarray:
        return self.rng[i * L.RNG_WORDS:(i + 1) * L.RNG_WORDS].cpu().numpy().view(np.uint32).copy()

    def set_rng_words(self, i: int, words: np.ndarray) -> None:
        w = torch.from_numpy(np.ascontiguousarray(words, dtype=np.uint32).view(np.int32))
        self.rng[i * L.RNG_WORDS:(i + 1) * L.RNG_WORDS].copy_(w.to(self.device))

    # -- runs ----------------------------------------------------------------
    def set_mu_adjust(self, values: Sequence[float]) -> None:
        for i, v in enumerate(values):
            self._tab_host[i].mu_adjust = float(v)
        self._upload_table()

    def run(self, num_steps: int, *, sync: bool = True, waves: Optional[int] = None) -> None:
        """mc_sim for every replica in one persistent launch, with ``waves``
        wavefronts per replica (default: ``self.waves``).  The result does not
        depend on ``waves``."""
        w = int(self.waves if waves is None else waves)
        if self.twist:
            w = 1       # the speculative evaluator has no twist terms
        L.check(self.lib.chromo_mc_run_tuned(self.tab_ptr, self.n, int(num_steps), w, self.binder_mask,
                                             _current_stream()), "chromo_mc_run_tuned")
        if sync:
            self.raise_on_error()

    def capture(self, num_steps: int = 1, *, waves: Optional[int] = None) -> "torch.cuda.CUDAGraph":
        """A HIP graph of ``run(num_steps)``: ``g.replay()`` advances every replica by
        num_steps sweeps with one graph launch (the launches record no memsets)."""
        w = 1 if self.twist else int(self.waves if waves is None else waves)
        self.run(0, waves=w)             # sets the launch attributes outside the capture
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream(self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                L.check(self.lib.chromo_mc_run_tuned(self.tab_ptr, self.n, int(num_steps), w,
                                                     self.binder_mask, _current_stream()),
                        "chromo_mc_run_tuned (capture)")
        torch.cuda.current_stream(self.device).wait_stream(s)
        return g

    def run_plain(self, num_steps: int, rep: int = 0, *, sync: bool = True) -> None:
        """The same schedule with one kernel launch per proposal."""
        L.check(self.lib.chromo_mc_run_plain(self.tab_ptr, rep, int(num_steps),
                                             _current_stream()), "chromo_mc_run_plain")
        if sync:
            self.raise_on_error()

    def step(self, ctrl: int, rep: int = 0) -> None:
        """One mc_step of controller ``ctrl``."""
        L.check(self.lib.chromo_mc_step(self.tab_ptr, rep, ctrl, _current_stream()),
                "chromo_mc_step")

    def update_amplitudes(self, ctrl: int, rep: int = 0) -> None:
        L.check(self.lib.chromo_update_amplitudes(self.tab_ptr, rep, ctrl, _current_stream()),
                "chromo_update_amplitudes")

    def raise_on_error(self) -> None:
        torch.cuda.synchronize(self.device)
        cnt = self.counters.view(self.n, L.CNT_WORDS)[:, L.CNT_ERROR].cpu().numpy()
        bad = np.nonzero(cnt)[0]
        if bad.size:
            code = int(cnt[bad[0]])
            raise L.ChromoError(f"replica {int(bad[0])}: {L.ERROR_TEXT.get(code, code)}")

    # -- checkpoint ----------------------------------------------------------
    def checkpoint(self, path) -> None:
        """Every replica's state (chains, states, densities, RNG, controllers,
        counters) to ``path``; ``restore`` continues it bit for bit."""
        L.check(self.lib.chromo_checkpoint(self.tab_ptr, self.n, str(path).encode(), _current_stream()),
                "chromo_checkpoint")

    def restore(self, path) -> None:
        L.check(self.lib.chromo_restore(self.tab_ptr, self.n, str(path).encode(), _current_stream()),
                "chromo_restore")

    # -- readback ------------------------------------------------------------
    def positions(self, i: int):
        p = self.pos[int(self.o_pos[i]):int(self.o_pos[i + 1])].view(9, self.Ns[i]).cpu().numpy()
        return p[0:3].T.copy(), p[3:6].T.copy(), p[6:9].T.copy()

    def set_positions(self, i: int, r, t3, t2) -> None:
        p = np.empty((9, self.Ns[i]))
        p[0:3] = np.asarray(r).T
        p[3:6] = np.asarray(t3).T
        p[6:9] = np.asarray(t2).T
        self.pos[int(self.o_pos[i]):int(self.o_pos[i + 1])].copy_(
            torch.from_numpy(p.reshape(-1)).to(self.device))

    def binder_states(self, i: int) -> np.ndarray:
        s = self.states[int(self.o_st[i]):int(self.o_st[i + 1])].view(self.Bs[i], self.Ns[i])
        return s.cpu().numpy().T.astype(np.int64)

    def set_binder_states(self, i: int, states) -> None:
        s = np.ascontiguousarray(np.asarray(states).T.reshape(-1).astype(np.int8))
        self.states[int(self.o_st[i]):int(self.o_st[i + 1])].copy_(torch.from_numpy(s).to(self.device))

    def density_of(self, i: int) -> np.ndarray:
        d = self.density[int(self.o_den[i]):int(self.o_den[i + 1])]
        return d.view(self.nbins[i], self.Bs[i] + 1).cpu().numpy().copy()

    def density_init(self) -> None:
        L.check(self.lib.chromo_density_init(self.tab_ptr, self.n, _current_stream()),
                "chromo_density_init")

    def controllers(self, i: int) -> List[CtrlSpec]:
        cs = ctypes.sizeof(L.ChromoCtrl)
        k = self.nctrl[i]
        a = cs * int(self.o_ctrl[i])
        raw = self.ctrl[a:a + cs * k].cpu().numpy().tobytes()
        arr = (L.ChromoCtrl * k).from_buffer_copy(raw)
        return [CtrlSpec.from_struct(c) for c in arr]

    def set_controllers(self, i: int, ctrls: Sequence[CtrlSpec]) -> None:
        """Replace replica i's controllers (schedule order); grows its arena if needed."""
        cs = ctypes.sizeof(L.ChromoCtrl)
        if len(ctrls) > self.CTRL_SLOTS:
            raise NotImplementedError(f"at most {self.CTRL_SLOTS} controllers per replica")
        bc, sc = default_caps(ctrls, self.Ns[i])
        if bc > 32 and (bc > self.bead_cap[i] or sc > self.seg_cap[i] or
                        not self._tab_host[i].scratch):
            if self.n != 1:
                raise NotImplementedError("growing the arena of a batched replica")
            self._grow_scratch(max(bc, self.bead_cap[i]), max(sc, self.seg_cap[i]))
        arr = (L.ChromoCtrl * len(ctrls))(*[c.to_struct() for c in ctrls])
        a = cs * int(self.o_ctrl[i])
        if len(ctrls):
            self.ctrl[a:a + cs * len(ctrls)].copy_(
                torch.frombuffer(bytearray(arr), dtype=torch.uint8).to(self.device))
        self.nctrl[i] = len(ctrls)
        self._tab_host[i].n_ctrl = len(ctrls)
        self._upload_table()

    def _upload_table(self) -> None:
        self.table.copy_(torch.frombuffer(bytearray(self._tab_host), dtype=torch.uint8).to(self.device))

    def _grow_scratch(self, bead_cap: int, seg_cap: int) -> None:
        nbytes = int(self.lib.chromo_scratch_bytes(bead_cap, seg_cap, self.Bs[0]))
        if nbytes < 0:
            raise ValueError("invalid arena capacities")
        torch.cuda.synchronize(self.device)
        self.scratch = torch.zeros(max(nbytes, 1), dtype=torch.uint8, device=self.device)
        self.bead_cap[0], self.seg_cap[0] = bead_cap, seg_cap
        R = self._tab_host[0]
        R.scratch = self.scratch.data_ptr()
        R.bead_cap, R.seg_cap = bead_cap, seg_cap
        self._upload_table()
        L.check(self.lib.chromo_scratch_init(self.tab_ptr, self.n, _current_stream()),
                "chromo_scratch_init")

    def counter_array(self) -> np.ndarray:
        return self.counters.view(self.n, L.CNT_WORDS).cpu().numpy().copy()

    def reset_counters(self) -> None:
        self.counters.zero_()

    def logs(self, i: int) -> np.ndarray:
        """Structured array of the per-proposal records of replica i."""
        if self.log is None:
            return np.zeros(0)
        ls = ctypes.sizeof(L.ChromoMoveLog)
        k = int(self.counter_array()[i, L.CNT_LOGGED])
        raw = self.log[ls * self.log_cap * i: ls * (self.log_cap * i + k)].cpu().numpy().tobytes()
        dt = np.dtype([("move", "<i4"), ("accepted", "<i4"), ("n", "<i8"), ("ind0", "<i8"),
                       ("indf", "<i8"), ("dE_poly", "<f8"), ("dE_field", "<f8"), ("u", "<f8"),
                       ("exp_dE", "<f8"), ("n_bins", "<i8")])
        return np.frombuffer(raw, dtype=dt).copy()

    # -- tape ------------------------------------------------------------------
    def field_dE_tape(self, i: int, inds, r_trial, states_trial, state_change: bool):
        """Field dE for given trial coordinates / states (no state change)."""
        dev = self.device
        inds = np.asarray(inds, dtype=np.int64).reshape(-1)
        n = int(inds.size)
        if n < 1 or inds.min() < 0 or inds.max() >= self.Ns[i]:
            raise ValueError("bead indices out of range")
        if np.asarray(r_trial).shape != (n, 3) or np.asarray(states_trial).shape != (n, self.Bs[i]):
            raise ValueError("r_trial must be (n, 3) and states_trial (n, B)")
        inds_t = torch.as_tensor(inds, device=dev)
        rt = torch.as_tensor(np.ascontiguousarray(r_trial, dtype=np.float64), device=dev)
        stt = torch.as_tensor(np.ascontiguousarray(states_trial, dtype=np.int8), device=dev)
        dE = torch.zeros(1, dtype=torch.float64, device=dev)
        bins = torch.zeros(16 * n, dtype=torch.int64, device=dev)
        delta = torch.zeros(16 * n * (self.Bs[i] + 1), dtype=torch.float64, device=dev)
        nb = torch.zeros(1, dtype=torch.int64, device=dev)
        L.check(self.lib.chromo_field_dE_tape(self.tab_ptr, i, inds_t.data_ptr(), n, rt.data_ptr(),
                                              stt.data_ptr(), int(state_change), dE.data_ptr(),
                                              bins.data_ptr(), delta.data_ptr(), nb.data_ptr(),
                                              _current_stream()), "chromo_field_dE_tape")
        torch.cuda.synchronize(dev)
        k = int(nb.item())
        return (float(dE.item()), bins[:k].cpu().numpy(),
                delta[:k * (self.Bs[i] + 1)].view(k, self.Bs[i] + 1).cpu().numpy())
