"""Python face of the SOFI2D time loop on MI355X.

``Problem`` reads a reference parameter file (and the model/source/receiver
files it names) through libsofi_host; ``Simulation`` runs shots on the GPU
through the C-ABI of libsofi_hip (include/sofi_hip.h).  ``run_json`` is the
Python counterpart of the ``sofi2D <file.json>`` driver (src/sofi2D.c:54-442
for MODE=FW): it writes the same SU seismograms, source signal and merged
snapshots.
"""
from __future__ import annotations

import ctypes as C
import os
import time

import numpy as np

from . import _native as N


class SofiError(RuntimeError):
    pass


def _chk_host(rc):
    if rc != 0:
        raise SofiError(N.host_error())


def _chk_hip(rc):
    if rc != 0:
        raise SofiError(N.hip_error())


def _as_int_ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _as_float_ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Problem:
    """A run set up from a JSON parameter file (params, model, acquisition)."""

    def __init__(self, json_path: str, base_dir: str | None = None):
        """base_dir: the directory the parameter file's relative file names
        (its own name included) are taken from; None: the working directory."""
        self._L = N.host()
        if base_dir:
            p = self._L.sofi_problem_load_in(json_path.encode(), os.fspath(base_dir).encode())
        else:
            p = self._L.sofi_problem_load(json_path.encode())
        if not p:
            raise SofiError(N.host_error())
        self._p = p
        self.nshots = self._L.sofi_problem_nshots(p)
        self.set_shot(1)

    def close(self):
        if self._p:
            self._L.sofi_problem_free(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ shots
    def set_shot(self, ishot: int):
        _chk_host(self._L.sofi_problem_set_shot(self._p, ishot))
        self.ishot = ishot
        self.desc = self._L.sofi_problem_desc(self._p).contents.copy()

    @property
    def nx(self):
        return self.desc.nx

    @property
    def ny(self):
        return self.desc.ny

    @property
    def plane(self):
        return N.pitch(self.desc.nx) * N.rows(self.desc.ny)

    def array(self, aid: int, planes: int = 1):
        ptr = self._L.sofi_problem_array(self._p, aid)
        if not ptr:
            return None
        n = self.plane * planes
        return np.ctypeslib.as_array(ptr, shape=(n,)).copy()

    def arrays_ptr(self):
        return self._L.sofi_problem_arrays(self._p)

    def profiles_ptr(self):
        return self._L.sofi_problem_profiles(self._p)

    def sources(self):
        n = self._L.sofi_problem_nsrc(self._p)
        L, p = self._L, self._p
        def ints(f):
            return np.ctypeslib.as_array(f(p), shape=(max(n, 1),))[:n].astype(np.int32).copy()
        def flts(f, m=1):
            return np.ctypeslib.as_array(f(p), shape=(max(n, 1) * m,))[: n * m].astype(np.float32).copy()
        return {
            "n": n, "i": ints(L.sofi_problem_src_i), "j": ints(L.sofi_problem_src_j),
            "type": ints(L.sofi_problem_src_type), "azi": flts(L.sofi_problem_src_azi),
            "amp": flts(L.sofi_problem_src_amp), "signals": flts(L.sofi_problem_signals, self.desc.nt),
            "source_type": L.sofi_problem_source_type(p),
        }

    def receivers(self):
        n = self._L.sofi_problem_ntr(self._p)
        ri = np.ctypeslib.as_array(self._L.sofi_problem_rec_i(self._p), shape=(max(n, 1),))[:n].copy()
        rj = np.ctypeslib.as_array(self._L.sofi_problem_rec_j(self._p), shape=(max(n, 1),))[:n].copy()
        return ri.astype(np.int32), rj.astype(np.int32)

    @property
    def snapcheck(self):
        return self._L.sofi_problem_snapcheck(self._p)

    @property
    def nsnap(self):
        return self._L.sofi_problem_nsnap(self._p)

    def seismo_comps(self):
        buf = (C.c_int * 5)()
        n = self._L.sofi_problem_seismo_comps(self._p, buf)
        return list(buf[:n])

    def snap_comps(self):
        buf = (C.c_int * 5)()
        n = self._L.sofi_problem_snap_comps(self._p, buf)
        return list(buf[:n])

    def snapshot_dims(self):
        d = self.desc
        nxs = (d.snap_idx[2] - d.snap_idx[1]) // max(d.idx, 1) + 1
        nys = (d.snap_idx[4] - d.snap_idx[3]) // max(d.idy, 1) + 1
        return nxs, nys

    # ------------------------------------------------------------ FWI
    @property
    def itermax(self) -> int:
        return self._L.sofi_problem_itermax(self._p)

    def model(self, which: int) -> np.ndarray:
        """The model as read, padded plane: 0 vp, 1 vs, 2 rho (MODE=FWI only)."""
        ptr = self._L.sofi_problem_model(self._p, which)
        if not ptr:
            raise SofiError("model planes are kept only for MODE=FWI")
        return np.ctypeslib.as_array(ptr, shape=(self.plane,)).copy()

    def read_data(self, comp: int, ntr: int) -> np.ndarray:
        """Measured data DATA_DIR_<comp>.su.shot<ishot> as [ntr][ns] (inseis)."""
        out = np.zeros((ntr, self.desc.ns), dtype=np.float32)
        _chk_host(self._L.sofi_read_data(self._L.sofi_problem_params(self._p), self.ishot, comp, ntr, _as_float_ptr(out)))
        return out

    def residual(self, data: np.ndarray, calc: np.ndarray):
        """Time-reversed L2 adjoint sources (calc_res) and the sum of their squares."""
        d = np.ascontiguousarray(data, dtype=np.float32)
        c = np.ascontiguousarray(calc, dtype=np.float32)
        out = np.zeros_like(c)
        l2 = C.c_double()
        _chk_host(self._L.sofi_residual(self._L.sofi_problem_params(self._p), c.shape[0], _as_float_ptr(d),
                                        _as_float_ptr(c), _as_float_ptr(out), C.byref(l2)))
        return out, l2.value

    def residual_opts(self) -> N.ResidualOpts:
        """LNORM, TIME_FILT (corners, ORDER) and WATERLEVEL_LNORM8 as the
        device residual takes them."""
        iv = (C.c_int * 3)()
        fv = (C.c_float * 3)()
        _chk_host(self._L.sofi_problem_residual_opts(self._p, iv, fv))
        return N.ResidualOpts(lnorm=iv[0], time_filt=iv[1], f_low_pass=fv[0], f_high_pass=fv[1], order=iv[2],
                              waterlevel_lnorm8=fv[2])

    def write_gradient(self, ishot: int, gvp: np.ndarray, gvs: np.ndarray, grho: np.ndarray):
        a = [np.ascontiguousarray(x, dtype=np.float32).ravel() for x in (gvp, gvs, grho)]
        _chk_host(self._L.sofi_write_gradient(self._L.sofi_problem_params(self._p), ishot, *[_as_float_ptr(x) for x in a]))

    # ------------------------------------------------------------ outputs
    def write_section(self, comp: int, section: np.ndarray):
        s = np.ascontiguousarray(section, dtype=np.float32)
        _chk_host(self._L.sofi_problem_write_section(self._p, comp, _as_float_ptr(s)))

    def write_signal(self):
        _chk_host(self._L.sofi_problem_write_signal(self._p))

    def write_snapshot(self, comp: int, first: bool, data: np.ndarray):
        s = np.ascontiguousarray(data, dtype=np.float32).ravel()
        _chk_host(self._L.sofi_problem_write_snapshot(self._p, comp, int(first), _as_float_ptr(s), s.size))


class Simulation:
    """Device state of one shot (a ``sofi_dev`` handle) for a Problem's grid."""

    def __init__(self, problem: Problem, desc: N.SofiDesc | None = None):
        self._L = N.hip()
        self.problem = problem
        self.desc = (desc or problem.desc).copy()
        h = C.c_void_p()
        _chk_hip(self._L.sofi_open(C.byref(self.desc), C.byref(h)))
        self._h = h
        _chk_hip(self._L.sofi_set_medium(h, problem.arrays_ptr(), problem.profiles_ptr()))
        ri, rj = problem.receivers()
        self.ntr = len(ri)
        self._ri, self._rj = ri, rj
        if self.desc.seismo and self.ntr:
            _chk_hip(self._L.sofi_set_receivers(h, self.ntr, _as_int_ptr(ri), _as_int_ptr(rj)))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.sofi_close(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def footprint(self) -> int:
        return int(self._L.sofi_footprint(self._h))

    def load_shot(self, nsnap: int | None = None):
        """Reset the wavefield and load the problem's current shot."""
        p = self.problem
        s = p.sources()
        self._src = s
        _chk_hip(self._L.sofi_reset(self._h))
        _chk_hip(self._L.sofi_set_sources(self._h, s["n"], _as_int_ptr(s["i"]), _as_int_ptr(s["j"]),
                                          _as_int_ptr(s["type"]), _as_float_ptr(s["azi"]),
                                          _as_float_ptr(s["amp"]), _as_float_ptr(s["signals"]),
                                          s["source_type"]))
        if nsnap is None:
            nsnap = p.nsnap if p.snapcheck else 0
        _chk_hip(self._L.sofi_set_snapshots(self._h, nsnap))

    def reset(self):
        """Zero the wavefield (zero_wavefield) without touching the sources."""
        _chk_hip(self._L.sofi_reset(self._h))

    # FWI (handle opened from a MODE=FWI problem)
    def fwi_set_model(self, vp0: np.ndarray, vs0: np.ndarray, rho0: np.ndarray):
        self._fwi_model = [np.ascontiguousarray(x, dtype=np.float32).ravel() for x in (vp0, vs0, rho0)]
        _chk_hip(self._L.sofi_fwi_set_model(self._h, *[_as_float_ptr(x) for x in self._fwi_model]))

    def fwi_set_pass(self, sw: int):
        _chk_hip(self._L.sofi_fwi_set_pass(self._h, sw))

    def fwi_set_adjoint(self, ri: np.ndarray, rj: np.ndarray, sig_x: np.ndarray, sig_y: np.ndarray):
        ri = np.ascontiguousarray(ri, dtype=np.int32)
        rj = np.ascontiguousarray(rj, dtype=np.int32)
        sx = np.ascontiguousarray(sig_x, dtype=np.float32)
        sy = np.ascontiguousarray(sig_y, dtype=np.float32)
        assert sx.shape == sy.shape == (len(ri), self.desc.nt), (sx.shape, len(ri), self.desc.nt)
        _chk_hip(self._L.sofi_fwi_set_adjoint(self._h, len(ri), _as_int_ptr(ri), _as_int_ptr(rj), _as_float_ptr(sx),
                                              _as_float_ptr(sy)))

    def fwi_residual(self, obs_x, obs_y, opts: N.ResidualOpts | None = None) -> float:
        """calc_res on the device: the adjoint sources at the receivers from
        the resident synthetic sections of the last forward pass and the
        measured sections ([ntr][ns], None for a component ADJOINT_TYPE does
        not use); returns the misfit."""
        opts = opts or self.problem.residual_opts()
        keep = [None if a is None else np.ascontiguousarray(a, dtype=np.float32) for a in (obs_x, obs_y)]
        for a in keep:
            assert a is None or a.shape == (self.ntr, self.desc.ns), (None if a is None else a.shape)
        ptr = [C.cast(None, C.POINTER(C.c_float)) if a is None else _as_float_ptr(a) for a in keep]
        mis = C.c_double()
        _chk_hip(self._L.sofi_fwi_residual(self._h, C.byref(opts), ptr[0], ptr[1], C.byref(mis)))
        return mis.value

    def fwi_adjoint(self, comp: int) -> np.ndarray:
        """The adjoint sources the next adjoint pass injects, [n][NT]."""
        n = int(self._L.sofi_fwi_adjoint_count(self._h))
        out = np.zeros((n, self.desc.nt), dtype=np.float32)
        if out.size:
            _chk_hip(self._L.sofi_fwi_get_adjoint(self._h, comp, _as_float_ptr(out)))
        return out

    def fwi_gradient(self):
        """The shot's (vp, vs, rho') gradient planes; adds them to the sums."""
        out = [np.zeros(self.problem.plane, dtype=np.float32) for _ in range(3)]
        _chk_hip(self._L.sofi_fwi_gradient(self._h, *[_as_float_ptr(x) for x in out]))
        return out

    def fwi_gradient_sum(self, reset: bool = False):
        out = [np.zeros(self.problem.plane, dtype=np.float32) for _ in range(3)]
        _chk_hip(self._L.sofi_fwi_gradient_sum(self._h, *[_as_float_ptr(x) for x in out], int(reset)))
        return out

    def run(self, nt_first: int = 1, nt_last: int | None = None) -> float:
        nt_last = self.desc.nt if nt_last is None else nt_last
        _chk_hip(self._L.sofi_run(self._h, nt_first, nt_last))
        return float(self._L.sofi_last_run_ms(self._h))

    LAUNCH = {"auto": 0, "cell": 1, "vec2": 2, "vec4": 3}

    def set_launch(self, shape: str, stress_shape: str | None = None):
        """Shapes of the isotropic/VTI update kernels: "auto", "cell" (one
        thread per cell), "vec2" or "vec4" (one thread per two or four cells
        of a row); `shape` for the velocity kernel and, unless
        `stress_shape` is given, the stress kernel."""
        st = shape if stress_shape is None else stress_shape
        _chk_hip(self._L.sofi_set_launch(self._h, self.LAUNCH[shape], self.LAUNCH[st]))

    GRAPH = {"auto": 0, "off": 1, "on": 2}

    def set_graph(self, mode: str = "auto", steps: int = 0):
        """Captured blocks of `steps` time steps (HIP graphs): "auto" (grids
        of at most 2^20 cells), "off" or "on"."""
        _chk_hip(self._L.sofi_set_graph(self._h, self.GRAPH[mode], steps))

    @property
    def graph_blocks(self) -> int:
        return int(self._L.sofi_graph_blocks(self._h))

    def set_kernel_timing(self, every: int, max_samples: int = 4096):
        _chk_hip(self._L.sofi_set_kernel_timing(self._h, every, max_samples))

    def kernel_timing(self):
        v, s, n = C.c_double(), C.c_double(), C.c_int()
        _chk_hip(self._L.sofi_get_kernel_timing(self._h, C.byref(v), C.byref(s), C.byref(n)))
        return {"velocity_ms": v.value, "stress_ms": s.value, "samples": n.value}

    def section(self, comp: int) -> np.ndarray:
        out = np.zeros((self.ntr, self.desc.ns), dtype=np.float32)
        if out.size:
            _chk_hip(self._L.sofi_get_section(self._h, comp, _as_float_ptr(out)))
        return out

    def snapshots(self, comp: int) -> np.ndarray:
        k = self._L.sofi_snapshot_count(self._h)
        nxs, nys = self.problem.snapshot_dims()
        out = np.zeros((k, nxs, nys), dtype=np.float32)
        for q in range(k):
            _chk_hip(self._L.sofi_get_snapshot(self._h, q, comp, _as_float_ptr(out[q])))
        return out

    def field(self, fid: int) -> np.ndarray:
        out = np.zeros(self.problem.plane, dtype=np.float32)
        _chk_hip(self._L.sofi_get_field(self._h, fid, _as_float_ptr(out)))
        return out.reshape(N.rows(self.desc.ny), N.pitch(self.desc.nx))


def run_many(sims, nt_first=None, nt_last=None, launch=None) -> float:
    """Advance one-shot simulations on grids of any size together
    (sofi_run_many): sims[i] through steps nt_first[i]..nt_last[i] (default
    1..NT of each), each step's velocity and stress updates of all of them in
    one work-list launch each.  launch = (rows, blocks, waves) fixes the
    launch; None lets the library measure it.  Returns the run's device time
    in ms."""
    n = len(sims)
    first = [1] * n if nt_first is None else list(nt_first)
    last = [s.desc.nt for s in sims] if nt_last is None else list(nt_last)
    L = N.hip()
    hs = (C.c_void_p * n)(*[s._h.value for s in sims])
    lp = (C.c_int * 3)(*launch) if launch is not None else None
    _chk_hip(L.sofi_run_many(hs, n, (C.c_int * n)(*first), (C.c_int * n)(*last), lp))
    return float(L.sofi_last_run_ms(sims[0]._h))


def many_config():
    """(rows, blocks, waves, tuned) of this thread's last run_many."""
    r, b, w, t = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    _chk_hip(N.hip().sofi_get_many_config(C.byref(r), C.byref(b), C.byref(w), C.byref(t)))
    return r.value, b.value, w.value, bool(t.value)


class Batch(Simulation):
    """Several shots of one Problem advanced together on one medium
    (sofi_open_batch): slot k holds shot shots[k]."""

    def __init__(self, problem: Problem, nslots: int):
        self._L = N.hip()
        self.problem = problem
        self.desc = problem.desc.copy()
        self.nslots = nslots
        h = C.c_void_p()
        _chk_hip(self._L.sofi_open_batch(C.byref(self.desc), nslots, C.byref(h)))
        self._h = h
        _chk_hip(self._L.sofi_set_medium(h, problem.arrays_ptr(), problem.profiles_ptr()))
        self.ntr = 0
        self._keep = []

    def load_shots(self, shots):
        """Sources and receivers of the problem's shots into slots 0.. and a
        zero wavefield in every slot."""
        assert len(shots) <= self.nslots
        p = self.problem
        self._keep = []
        for k, ishot in enumerate(shots):
            p.set_shot(ishot)
            _chk_hip(self._L.sofi_select_shot(self._h, k))
            s = p.sources()
            self._keep.append(s)
            _chk_hip(self._L.sofi_set_sources(self._h, s["n"], _as_int_ptr(s["i"]), _as_int_ptr(s["j"]),
                                              _as_int_ptr(s["type"]), _as_float_ptr(s["azi"]),
                                              _as_float_ptr(s["amp"]), _as_float_ptr(s["signals"]),
                                              s["source_type"]))
            ri, rj = p.receivers()
            self.ntr = len(ri)
            if self.desc.seismo and self.ntr:
                _chk_hip(self._L.sofi_set_receivers(self._h, self.ntr, _as_int_ptr(ri), _as_int_ptr(rj)))
        _chk_hip(self._L.sofi_select_shot(self._h, 0))
        _chk_hip(self._L.sofi_reset(self._h))

    def select(self, k: int):
        _chk_hip(self._L.sofi_select_shot(self._h, k))

    def set_launch_groups(self, group: int = 0, spt: int = 0):
        """Slots per launch and per thread of the update kernels (0: tuned)."""
        _chk_hip(self._L.sofi_set_batch_launch(self._h, group, spt))

    def set_resident(self, slots: int = 0):
        """Slots advanced through a run's whole step range before the next
        group (0: the largest group that fits the memory-side cache)."""
        _chk_hip(self._L.sofi_set_batch_resident(self._h, slots))

    def set_block(self, slots: int = 0):
        """Slot blocks: all slots in one launch per kernel, `slots` (1, 2 or
        4) slots per thread; 0 for the slot-loop kernels."""
        _chk_hip(self._L.sofi_set_batch_block(self._h, slots))

    @property
    def block(self) -> int:
        return int(self._L.sofi_get_batch_block(self._h))

    def set_tuning(self, on: bool = True):
        """The launch autotuner (on by default; fixing any knob turns it off)."""
        _chk_hip(self._L.sofi_set_batch_tuning(self._h, int(on)))

    @property
    def config(self) -> dict:
        """The launch configuration in use (chosen by the autotuner unless fixed)."""
        c = (C.c_int * 6)()
        _chk_hip(self._L.sofi_get_batch_config(self._h, c))
        return {"slot_block": c[0], "resident": c[1], "streams": c[2], "slots_per_launch": c[3],
                "slots_per_thread": c[4], "tuned": bool(c[5])}

    def set_streams(self, streams: int = 0):
        """Streams the resident groups run on at once (0: 2)."""
        _chk_hip(self._L.sofi_set_batch_streams(self._h, streams))

    @property
    def resident(self) -> int:
        return int(self._L.sofi_get_batch_resident(self._h))

    @property
    def launch_groups(self):
        g, s = C.c_int(), C.c_int()
        _chk_hip(self._L.sofi_get_batch_launch(self._h, C.byref(g), C.byref(s)))
        return g.value, s.value

    def section_of(self, k: int, comp: int) -> np.ndarray:
        self.select(k)
        return self.section(comp)


def interior(a: np.ndarray, nx: int, ny: int) -> np.ndarray:
    """[ny][nx] view of the interior of a padded-layout plane."""
    a = a.reshape(N.rows(ny), N.pitch(nx))
    return a[N.PAD:N.PAD + ny, N.PAD:N.PAD + nx]


def run_json(json_path: str, verbose: bool = True) -> dict:
    """sofi2D <file.json> on the GPU: every shot, all output files."""
    t0 = time.time()
    prob = Problem(json_path)
    sim = Simulation(prob)
    loop_s = 0.0
    snaps_written = 0
    results = {}
    for ishot in range(1, prob.nshots + 1):
        if ishot > 1:
            prob.set_shot(ishot)
        prob.write_signal()
        sim.load_shot()
        ms = sim.run()
        loop_s += ms * 1e-3
        if verbose:
            d = prob.desc
            print(f"shot {ishot}: time loop {ms * 1e-3:.3f} s ({d.nx * d.ny * d.nt / (ms * 1e-3) / 1e9:.2f} GLUPS)")
        secs = {}
        for c in prob.seismo_comps():
            secs[c] = sim.section(c)
            prob.write_section(c, secs[c])
        snaps = {}
        comps = prob.snap_comps()
        for c in comps:
            snaps[c] = sim.snapshots(c)
        if comps and len(snaps[comps[0]]):
            for k in range(len(snaps[comps[0]])):
                for c in comps:
                    prob.write_snapshot(c, snaps_written == 0, snaps[c][k])
                snaps_written += 1
        results[ishot] = {"sections": secs, "snapshots": snaps, "ms": ms}
    if verbose:
        print(f"total time loop {loop_s:.3f} s, program {time.time() - t0:.3f} s")
    sim.close()
    prob.close()
    return results
