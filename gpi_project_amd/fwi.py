"""MODE=FWI of sofi2D on the GPU: the gradient of the misfit for every shot
(src/sofi2D.c:280-404 with src/inversion.c:27-258).

Per iteration (ITERMAX), the gradient sums are cleared (init_inv,
src/init_inv.c:20-25); per shot:
  forward pass  time loop with the shot's sources, storing the forward
                fields every DTINV steps; the vx/vy seismograms are written
                as in MODE=FW (saveseis)
  residuals     measured data DATA_DIR_<vx|vy>.su.shot<n> (inseis), the
                adjoint sources of LNORM, time-reversed (calc_res), formed on
                the device from the resident synthetic sections
                (sofi_fwi_residual; host_residual=True forms them on the
                host from sections read back, as the reference's loop does)
  adjoint pass  the residuals injected at the receivers (ADJOINT_TYPE),
                zero-lag correlation with the stored forward fields
  gradient      vp / vs / rho' of the shot, written as
                GRADIENT_gradient.<p>.shot<n>; after the last shot the sums
                GRADIENT_gradient.<p> (output_gradient)
The model update (step length, L-BFGS) is not part of this path.
"""
from __future__ import annotations

import time

import numpy as np

from .sofi import Batch, Problem, Simulation

VX, VY = 0, 1


def shot_gradient(prob: Problem, sim: Simulation, ishot: int, write: bool = True, sections: bool = True,
                  host_residual: bool = False) -> dict:
    """Forward pass, residuals, adjoint pass and gradient of one shot.
    sections: also return the synthetic sections (read back from the
    device; they are read back anyway when written or for host residuals)."""
    d = prob.desc
    if ishot != prob.ishot:
        prob.set_shot(ishot)
    t0 = time.perf_counter()
    sim.fwi_set_pass(0)
    sim.load_shot(nsnap=prob.nsnap if prob.snapcheck else 0)
    ms_fwd = sim.run()
    secs = {}
    if write or sections or host_residual:
        secs = {c: sim.section(c) for c in prob.seismo_comps()}
    if write:
        for c, a in secs.items():
            prob.write_section(c, a)
    ri, rj = prob.receivers()
    ntr = len(ri)
    obs_x = prob.read_data(VX, ntr) if d.adjoint_type in (1, 3) else None
    obs_y = prob.read_data(VY, ntr) if d.adjoint_type in (1, 2) else None
    if host_residual:
        zero = np.zeros((ntr, d.ns), dtype=np.float32)
        diff_x, diff_y, l2 = zero, zero, 0.0
        if obs_x is not None:
            diff_x, e = prob.residual(obs_x, secs[VX])
            l2 += e
        if obs_y is not None:
            diff_y, e = prob.residual(obs_y, secs[VY])
            l2 += e
        # NS == NT (NDT = 1): the adjoint signal of step nt is sample nt
        sim.fwi_set_adjoint(ri, rj, diff_x, diff_y)
    else:
        l2 = sim.fwi_residual(obs_x, obs_y)
    sim.fwi_set_pass(1)
    sim.reset()
    ms_adj = sim.run()
    gvp, gvs, grho = sim.fwi_gradient()
    if write:
        prob.write_gradient(ishot, gvp, gvs, grho)
    return {"gradient": (gvp, gvs, grho), "sections": secs, "l2": l2, "ms_forward": ms_fwd, "ms_adjoint": ms_adj,
            "s": time.perf_counter() - t0}


def gradient_json(json_path: str, verbose: bool = True, write: bool = True, workers: int = 1,
                  sections: bool = True, host_residual: bool = False, batch: int = 1) -> dict:
    """Every iteration and shot of a MODE=FWI parameter file; returns the
    per-shot results of the last iteration and the summed gradients.

    workers > 1 computes that many shots at once, each on its own Problem,
    handle and HIP stream (one host thread each), so that the kernels of
    several shots overlap on the device; the shots' float32 gradients are
    then summed in shot order, the order of the one-handle path, which gives
    the same sums bit for bit.

    batch > 1 advances groups of that many shots as the slots of one batch
    handle sharing the medium (each update launch covers every slot of the
    group: forward passes together, residuals slot by slot, adjoint passes
    together); with workers > 1 as well, that many such handles run at once.
    The residuals are formed on the device."""
    if batch > 1:
        return _gradient_json_batched(json_path, verbose, write, workers, sections, batch)
    if workers > 1:
        return _gradient_json_concurrent(json_path, verbose, write, workers, sections, host_residual)
    prob = Problem(json_path)
    if prob.desc.mode != 0:
        raise ValueError(f"{json_path}: MODE is not FWI")
    sim = Simulation(prob)
    sim.fwi_set_model(prob.model(0), prob.model(1), prob.model(2))
    shots = {}
    for it in range(1, prob.itermax + 1):
        sim.fwi_gradient_sum(reset=True)
        for ishot in range(1, prob.nshots + 1):
            r = shot_gradient(prob, sim, ishot, write, sections, host_residual)
            shots[ishot] = r
            if verbose:
                print(f"iteration {it} shot {ishot}: forward {r['ms_forward']:.1f} ms, adjoint {r['ms_adjoint']:.1f} ms, "
                      f"L2 {r['l2']:.6e}")
    gsum = sim.fwi_gradient_sum()
    if write:
        prob.write_gradient(0, *gsum)
    sim.close()
    prob.close()
    return {"shots": shots, "sum": gsum}


def _gradient_json_concurrent(json_path: str, verbose: bool, write: bool, workers: int, sections: bool,
                              host_residual: bool) -> dict:
    import threading
    probe = Problem(json_path)
    if probe.desc.mode != 0:
        probe.close()
        raise ValueError(f"{json_path}: MODE is not FWI")
    nshots, itermax = probe.nshots, probe.itermax
    workers = min(workers, nshots)
    lock = threading.Lock()
    shots: dict = {}
    errors: list = []
    for it in range(1, itermax + 1):
        queue = list(range(1, nshots + 1))

        def worker():
            prob = sim = None
            try:
                prob = Problem(json_path)
                sim = Simulation(prob)
                sim.fwi_set_model(prob.model(0), prob.model(1), prob.model(2))
                while True:
                    with lock:
                        if not queue or errors:
                            return
                        ishot = queue.pop(0)
                    r = shot_gradient(prob, sim, ishot, write, sections, host_residual)
                    with lock:
                        shots[ishot] = r
                    if verbose:
                        print(f"iteration {it} shot {ishot}: forward {r['ms_forward']:.1f} ms, "
                              f"adjoint {r['ms_adjoint']:.1f} ms, L2 {r['l2']:.6e}")
            except Exception as e:  # noqa: BLE001
                with lock:
                    errors.append(e)
            finally:
                if sim is not None:
                    sim.close()
                if prob is not None:
                    prob.close()

        threads = [threading.Thread(target=worker) for _ in range(workers)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        if errors:
            probe.close()
            raise errors[0]
    gsum = [np.zeros(probe.plane, dtype=np.float32) for _ in range(3)]
    for ishot in range(1, nshots + 1):
        for q in range(3):
            gsum[q] = gsum[q] + shots[ishot]["gradient"][q]
    if write:
        probe.write_gradient(0, *gsum)
    probe.close()
    return {"shots": shots, "sum": gsum}


def _gradient_json_batched(json_path: str, verbose: bool, write: bool, workers: int, sections: bool,
                           batch: int) -> dict:
    import threading
    probe = Problem(json_path)
    if probe.desc.mode != 0:
        probe.close()
        raise ValueError(f"{json_path}: MODE is not FWI")
    nshots, itermax = probe.nshots, probe.itermax
    batch = min(batch, nshots)
    groups_all = [list(range(a, min(a + batch, nshots + 1))) for a in range(1, nshots + 1, batch)]
    workers = max(1, min(workers, len(groups_all)))
    lock = threading.Lock()
    shots: dict = {}
    errors: list = []
    for it in range(1, itermax + 1):
        queue = list(groups_all)

        def worker():
            prob = B = None
            try:
                prob = Problem(json_path)
                B = Batch(prob, batch)
                B.fwi_set_model(prob.model(0), prob.model(1), prob.model(2))
                d = prob.desc
                while True:
                    with lock:
                        if not queue or errors:
                            return
                        group = queue.pop(0)
                    t0 = time.perf_counter()
                    B.fwi_set_pass(0)
                    B.load_shots(group)
                    ms_fwd = B.run()
                    res = {}
                    for k, ishot in enumerate(group):
                        B.select(k)
                        prob.set_shot(ishot)
                        secs = {}
                        if write or sections:
                            secs = {c: B.section(c) for c in prob.seismo_comps()}
                        if write:
                            for c, a in secs.items():
                                prob.write_section(c, a)
                        obs_x = prob.read_data(VX, B.ntr) if d.adjoint_type in (1, 3) else None
                        obs_y = prob.read_data(VY, B.ntr) if d.adjoint_type in (1, 2) else None
                        res[ishot] = {"sections": secs, "l2": B.fwi_residual(obs_x, obs_y)}
                    B.fwi_set_pass(1)
                    B.select(0)
                    B.reset()
                    ms_adj = B.run()
                    for k, ishot in enumerate(group):
                        B.select(k)
                        g = B.fwi_gradient()
                        if write:
                            prob.write_gradient(ishot, *g)
                        res[ishot].update({"gradient": tuple(g), "ms_forward": ms_fwd / len(group),
                                           "ms_adjoint": ms_adj / len(group), "s": time.perf_counter() - t0})
                    with lock:
                        shots.update(res)
                    if verbose:
                        print(f"iteration {it} shots {group}: forward {ms_fwd:.1f} ms, adjoint {ms_adj:.1f} ms")
            except Exception as e:  # noqa: BLE001
                with lock:
                    errors.append(e)
            finally:
                if B is not None:
                    B.close()
                if prob is not None:
                    prob.close()

        threads = [threading.Thread(target=worker) for _ in range(workers)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        if errors:
            probe.close()
            raise errors[0]
    gsum = [np.zeros(probe.plane, dtype=np.float32) for _ in range(3)]
    for ishot in range(1, nshots + 1):
        for q in range(3):
            gsum[q] = gsum[q] + shots[ishot]["gradient"][q]
    if write:
        probe.write_gradient(0, *gsum)
    probe.close()
    return {"shots": shots, "sum": gsum}
