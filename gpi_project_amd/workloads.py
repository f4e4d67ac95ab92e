"""Synthetic inputs in the reference's file formats.

``write_case`` writes one run directory: the pseudo-JSON parameter file
(src/read_par_json.c keys), raw float32 x-major model files
MFILE.{vp,vs,rho[,qp,qs][,epsilon,delta][,theta]} (src/readmod_*.c), a
SOURCE_FILE ("x y td fc amp [azimuth] [type]" per line, src/sources.c:78-131)
and, when receivers are listed explicitly, a REC_FILE ("x y" per line,
src/receiver.c:68-74).  The ``config*`` functions build the benchmark
workloads with seeded generators (numpy PCG64).
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

WEQ = ("AC_ISO", "VAC_ISO", "EL_ISO", "VEL_ISO", "EL_VTI", "VEL_VTI", "EL_TTI", "VEL_TTI")


@dataclass
class Case:
    name: str
    nx: int
    ny: int
    dh: float
    dt: float
    time: float
    weq: str = "EL_ISO"
    fdorder: int = 4
    fdorder_time: int = 2
    maxrelerror: int = 0
    L: int = 0
    fl: tuple = (15.0,)
    f_ref: float = 15.0
    abs_type: int = 1
    fw: int = 20
    free_surf: int = 0
    damping: float = 8.0
    npower: float = 4.0
    k_max_cpml: float = 1.0
    fpml: float = 15.0
    vppml: float = 3500.0
    source_type: int = 1
    source_shape: int = 1
    sources: list = field(default_factory=list)    # tuples (x, y, td, fc, amp[, azimuth, type])
    receivers: list | None = None                  # explicit (x, y) list, or None for a line
    rec_line: tuple = (0.0, 0.0, 0.0, 0.0, 1.0)      # XREC1, YREC1, XREC2, YREC2, NGEOPH
    seismo: int = 1
    ndt: int = 1
    snap: int = 0
    tsnap: tuple = (0.0, 0.0, 0.0)
    idx: int = 1
    idy: int = 1
    run_multiple_shots: int = 0
    mode: str = "FW"
    extra: dict = field(default_factory=dict)


def _json_lines(c: Case, d: str) -> str:
    rel = lambda *p: "./" + "/".join(p)
    kv = [
        ("NPROCX", 1), ("NPROCY", 1), ("MODE", c.mode),
        ("FDORDER", c.fdorder), ("FDORDER_TIME", c.fdorder_time), ("MAXRELERROR", c.maxrelerror),
        ("NX", c.nx), ("NY", c.ny), ("DH", repr(float(c.dh))),
        ("TIME", repr(float(c.time))), ("DT", repr(float(c.dt))),
        ("WEQ", c.weq),
        ("SOURCE_SHAPE", c.source_shape), ("SOURCE_TYPE", c.source_type), ("SIGOUT", 0),
        ("SRCREC", 1), ("SOURCE_FILE", rel("sources.dat")), ("RUN_MULTIPLE_SHOTS", c.run_multiple_shots),
        ("READMOD", 1), ("MFILE", rel("model", "m")), ("WRITE_MODELFILES", 0),
        ("L", c.L), ("F_REF", repr(float(c.f_ref))),
        ("FREE_SURF", c.free_surf), ("BOUNDARY", 0), ("FW", c.fw), ("ABS_TYPE", c.abs_type),
        ("NPOWER", repr(float(c.npower))), ("K_MAX_CPML", repr(float(c.k_max_cpml))),
        ("VPPML", repr(float(c.vppml))), ("FPML", repr(float(c.fpml))), ("DAMPING", repr(float(c.damping))),
        ("SNAP", c.snap), ("TSNAP1", repr(float(c.tsnap[0]))), ("TSNAP2", repr(float(c.tsnap[1]))),
        ("TSNAPINC", repr(float(c.tsnap[2]))), ("IDX", c.idx), ("IDY", c.idy), ("SNAP_FORMAT", 3),
        ("SNAP_FILE", rel("out", "snap")),
        ("SEISMO", c.seismo), ("READREC", 1 if c.receivers is not None else 0),
        ("REC_FILE", rel("receivers.dat")),
        ("XREC1", repr(float(c.rec_line[0]))), ("YREC1", repr(float(c.rec_line[1]))),
        ("XREC2", repr(float(c.rec_line[2]))), ("YREC2", repr(float(c.rec_line[3]))),
        ("NGEOPH", repr(float(c.rec_line[4]))),
        ("REFRECX", "0.0"), ("REFRECY", "0.0"), ("REC_ARRAY", 0),
        ("NDT", c.ndt), ("SEIS_FORMAT", 1), ("SEIS_FILE", rel("out", "seis")),
        ("LOG_FILE", rel("out", "log")), ("LOG", 0), ("OUT_TIMESTEP_INFO", 1000000),
    ]
    for l, f in enumerate(c.fl[: c.L], start=1):
        kv.append((f"FL{l}", repr(float(f))))
    # extra keys replace the defaults above, or are added after them
    keys = [k for k, _ in kv]
    for k, v in c.extra.items():
        if k in keys:
            kv[keys.index(k)] = (k, v)
        else:
            kv.append((k, v))
    body = ",\n".join(f'\t"{k}" : "{v}"' for k, v in kv)
    return "{\n" + body + "\n}\n"


def write_case(d: str, c: Case, model: dict) -> str:
    """Write the case into directory d; model maps suffix -> [nx][ny] array."""
    os.makedirs(os.path.join(d, "model"), exist_ok=True)
    os.makedirs(os.path.join(d, "out"), exist_ok=True)
    for k, a in model.items():
        a = np.asarray(a, dtype=np.float32)
        assert a.shape == (c.nx, c.ny), (k, a.shape)
        a.tofile(os.path.join(d, "model", f"m.{k}"))      # x-major: NY floats per x position
    with open(os.path.join(d, "sources.dat"), "w") as f:
        for s in c.sources:
            f.write(" ".join(repr(float(v)) for v in s) + "\n")
    if c.receivers is not None:
        with open(os.path.join(d, "receivers.dat"), "w") as f:
            for x, y in c.receivers:
                f.write(f"{float(x)!r} {float(y)!r}\n")
    path = os.path.join(d, f"{c.name}.json")
    with open(path, "w") as f:
        f.write(_json_lines(c, d))
    return path


# ----------------------------------------------------------------- models

def smooth_noise(rng, nx, ny, sigma_cells):
    """White noise smoothed by a separable Gaussian (FFT), normalised to unit std."""
    w = rng.standard_normal((nx, ny))
    kx = np.fft.fftfreq(nx)[:, None]
    ky = np.fft.rfftfreq(ny)[None, :]
    g = np.exp(-2.0 * (math.pi * sigma_cells) ** 2 * (kx ** 2 + ky ** 2))
    s = np.fft.irfft2(np.fft.rfft2(w) * g, s=(nx, ny))
    s -= s.mean()
    return s / (s.std() + 1e-30)


def layered_model(nx, ny, dh, rng, nlayers=12, vp_top=1800.0, vp_bottom=4500.0, pert=0.05,
                  corr_m=50.0, vp_max=None):
    """Flat layers with vp linear in layer index, vs = vp/1.73, rho = 310 vp^0.25,
    plus a smoothed Gaussian perturbation of std `pert` clipped at 3 sigma."""
    depth = np.arange(ny) * dh
    layer = np.minimum((depth / (ny * dh) * nlayers).astype(int), nlayers - 1)
    vp1 = vp_top + (vp_bottom - vp_top) * layer / max(nlayers - 1, 1)
    vp = np.repeat(vp1[None, :], nx, axis=0)
    if pert > 0:
        r = np.clip(smooth_noise(rng, nx, ny, corr_m / dh), -3.0, 3.0)
        vp = vp * (1.0 + pert * r)
    if vp_max is not None:
        vp = np.minimum(vp, vp_max)
    vs = vp / 1.73
    rho = 310.0 * vp ** 0.25
    return vp.astype(np.float32), vs.astype(np.float32), rho.astype(np.float32)


def q_from_tau(tau, L):
    """Q giving the reference's tau = 2/(Q L) (src/readmod_visco.c:156-157)."""
    return 2.0 / (tau * L)


# ---------------------------------------------------------------- configs

def config1(d: str) -> str:
    """Isotropic elastic, 400 x 400, DH 5 m, FD4 Holberg, DT 5e-4 s, NT 2000,
    homogeneous vp 3500 / vs 2000 / rho 2000, 15 Hz Ricker explosive source at
    the centre, 20-point CPML, 100 receivers on a line 50 m below the top."""
    nx = ny = 400
    c = Case(name="c1", nx=nx, ny=ny, dh=5.0, dt=5e-4, time=1.0, weq="EL_ISO", fdorder=4, maxrelerror=1,
             abs_type=1, fw=20, fpml=15.0, vppml=3500.0,
             sources=[(997.5, 997.5, 0.0, 15.0, 1.0)], seismo=1, ndt=1,
             rec_line=(200.0, 50.0, 1685.0, 50.0, 3.0))
    model = {"vp": np.full((nx, ny), 3500.0), "vs": np.full((nx, ny), 2000.0), "rho": np.full((nx, ny), 2000.0)}
    return write_case(d, c, model)


C2_FL = (5.0, 15.0, 45.0)


def config2(d: str, nx=4000, ny=2000, nt=8000, seed=2, weq="VEL_ISO", fdorder=8, name="c2") -> str:
    """Viscoelastic L=3 (tau=0.05 at 15 Hz), FD8, DH 2.5 m, DT 2.5e-4 s, free
    surface + 20-point CPML; 12 flat layers vp 1800..4500 m/s with a 5 %
    smoothed Gaussian perturbation (50 m correlation length, seed 2),
    vs = vp/1.73, rho = 310 vp^0.25; 15 Hz Ricker explosive source at
    x = 5000 m, y = 25 m; 1000 receivers at y = 12.5 m every 2 cells from 250 m.
    The perturbed vp is clipped at 5060 m/s so the unrelaxed velocity
    (1.034 vp for these relaxation frequencies) stays below the FD8 stability
    limit checkfd enforces (DT <= DH / (sqrt(2) sum|hc| vmax), 5254 m/s).
    weq = "EL_ISO" drops the relaxation (no Q files); fdorder 4 is stable at
    the same DT (its Holberg sum|hc| is smaller)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    dh = 2.5
    vp, vs, rho = layered_model(nx, ny, dh, rng, vp_max=5060.0)
    visco = weq == "VEL_ISO"
    L, tau = (3, 0.05) if visco else (0, 0.0)
    x_src = min(5000.0, (nx - 1) * dh * 0.5)
    ntr = min(1000, max(1, (nx - 2 * 20 - 100) // 2))
    x1 = 250.0 if nx > 400 else 20 * dh
    c = Case(name=name, nx=nx, ny=ny, dh=dh, dt=2.5e-4, time=nt * 2.5e-4, weq=weq, fdorder=fdorder,
             maxrelerror=1, L=L, fl=C2_FL, f_ref=15.0, abs_type=1, fw=20, free_surf=1, fpml=15.0,
             vppml=float(vp.max()), sources=[(x_src, 25.0, 0.0, 15.0, 1.0)], seismo=1, ndt=4,
             rec_line=(x1, 12.5, x1 + 2 * dh * (ntr - 1), 12.5, 2.0))
    model = {"vp": vp, "vs": vs, "rho": rho}
    if visco:
        q = np.full((nx, ny), q_from_tau(tau, L), dtype=np.float32)
        model.update(qp=q, qs=q)
    return write_case(d, c, model)


@dataclass
class Request4:
    rid: int
    nx: int
    ny: int
    nt: int
    weq: str
    fdorder: int
    gap: float          # unit-rate exponential inter-arrival gap


def config4_requests(nreq: int, seed: int = 5, side=(256, 4096), nt=(1000, 8000)) -> list:
    """Served shot requests: grid sides log-uniform in [256, 4096]
    (independently in x and y), FDORDER in {4, 8}, NT ~ U{1000..8000}, WEQ in
    {EL_ISO, VEL_ISO}, and unit-rate exponential inter-arrival gaps (Poisson
    arrivals; multiply by the mean service time / offered load), all from one
    generator of seed 5.  Each request's model is the C2 model with seed
    100 + rid."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lo, hi = math.log(side[0]), math.log(side[1])
    out = []
    for r in range(nreq):
        nx = int(round(math.exp(rng.uniform(lo, hi))))
        ny = int(round(math.exp(rng.uniform(lo, hi))))
        fd = int(rng.choice([4, 8]))
        n = int(rng.integers(nt[0], nt[1] + 1))
        weq = ("EL_ISO", "VEL_ISO")[int(rng.integers(0, 2))]
        out.append(Request4(r, nx, ny, n, weq, fd, float(rng.exponential(1.0))))
    return out


def config4_write(d: str, q: "Request4") -> str:
    """Run directory of one served request (C2 model of seed 100 + rid)."""
    return config2(d, nx=q.nx, ny=q.ny, nt=q.nt, seed=100 + q.rid, weq=q.weq, fdorder=q.fdorder,
                   name=f"c4_{q.rid:03d}")


def _rank_uniform(field: np.ndarray, lo: float, hi: float) -> np.ndarray:
    """Map a field to U(lo, hi) by its ranks (keeps its smooth structure)."""
    r = np.empty(field.size)
    r[np.argsort(field, axis=None, kind="stable")] = np.arange(field.size)
    return (lo + (hi - lo) * (r + 0.5) / field.size).reshape(field.shape)


def config3(d: str, weq: str = "EL_VTI", nshots: int = 64, nx: int = 2000, ny: int = 1000, nt: int = 4000,
            seed: int = 3) -> str:
    """Batched shots: EL_VTI or EL_TTI (or EL_ISO), 2000 x 1000, DH 5 m,
    FD6 (Holberg 0.1 %), DT 4e-4 s, 20-point CPML, no free surface.  Shared
    model: 12 flat layers vp 1800..4500 m/s (vs = vp/1.73, rho = 310
    vp^0.25); epsilon ~ U(0, 0.25), delta ~ U(-0.1, 0.15), theta ~ U(0, 45)
    degrees as smooth fields (uniform noise of seed 3 smoothed with a 100 m
    Gaussian, rank-mapped to the ranges).  nshots shots (seed 4): x ~ U(FW
    DH, (NX - FW) DH), y = 25 m, 10..25 Hz Ricker explosive sources; 500
    receivers at y = 12.5 m every 3 cells from cell 250."""
    dh = 5.0
    rng = np.random.Generator(np.random.PCG64(seed))
    vp, vs, rho = layered_model(nx, ny, dh, rng, pert=0.0)
    model = {"vp": vp, "vs": vs, "rho": rho}
    sm = lambda: smooth_noise(rng, nx, ny, 100.0 / dh)
    if weq != "EL_ISO":
        model["epsilon"] = _rank_uniform(sm(), 0.0, 0.25).astype(np.float32)
        model["delta"] = _rank_uniform(sm(), -0.1, 0.15).astype(np.float32)
    if weq == "EL_TTI":
        model["theta"] = _rank_uniform(sm(), 0.0, 45.0).astype(np.float32)
    srng = np.random.Generator(np.random.PCG64(4))
    fw = 20
    xs = srng.uniform(fw * dh, (nx - fw) * dh, nshots)
    fcs = srng.uniform(10.0, 25.0, nshots)
    ntr = min(500, max(1, (nx - 2 * fw - 250) // 3))
    x1 = 250 * dh if nx > 2000 - 1 else (fw + 5) * dh
    c = Case(name=f"c3_{weq.lower()}", nx=nx, ny=ny, dh=dh, dt=4.0e-4, time=nt * 4.0e-4, weq=weq, fdorder=6,
             maxrelerror=1, abs_type=1, fw=fw, free_surf=0, fpml=15.0, vppml=float(vp.max()),
             sources=[(float(x), 25.0, 0.0, float(f), 1.0) for x, f in zip(xs, fcs)], seismo=1, ndt=4,
             rec_line=(x1, 12.5, x1 + 3 * dh * (ntr - 1), 12.5, 3.0), run_multiple_shots=1)
    return write_case(d, c, model)


FWI_KEYS = {"ITERMAX": 1, "PARAMETERIZATION": 1, "INV_VP_ITER": 1, "INV_VS_ITER": 1, "INV_RHO_ITER": 1,
            "TIME_FILT": 0, "F_HIGH_PASS": 1, "F_LOW_PASS_START": "10.0", "F_LOW_PASS_END": "75.0",
            "F_LOW_PASS_INCR": "10.0", "ORDER": 8, "ZERO_PHASE": 0, "FREQ_FILE": "frequencies.dat",
            "WRITE_FILTERED_DATA": 0, "MIN_ITER": 10, "USE_WORKFLOW": 0, "FILE_WORKFLOW": "workflow.txt",
            "SWS_TAPER_GRAD_SOURCES": 0, "SRTRADIUS": 0, "SRTSHAPE": 0, "FILTSIZE": 0,
            "INV_MODELFILE": "./grad/model", "NFSTART": 1, "NF": 1, "GRADIENT": "./grad/g", "NFSTART_GRAD": 1,
            "NF_GRAD": 1, "GRAD_METHOD": 2, "LNORM": 2, "NORMALIZE": 0, "WRITE_DIFF": 0, "EPS_SCALE": "0.01",
            "STEPMAX": 4, "SCALEFAC": "4.0", "TESTSHOT_START": 1, "TESTSHOT_END": 2, "TESTSHOT_INCR": 1,
            "PRO": "0.01", "MISFIT_LOG_FILE": "./grad/L2_LOG.dat"}


def config5(d: str, which: str, nshots: int = 32, nx: int = 1500, ny: int = 500, nt: int = 3000,
            dtinv: int = 5, amp: float = 1e9) -> str:
    """Elastic FWI: 1500 x 500, DH 5 m, FD8 (Holberg 0.1 %), DT 5e-4 s,
    NT 3000, 20-point CPML all round.  Background vp 2000 -> 4000 m/s linear
    in depth, vs = vp/1.73, rho = 310 vp^0.25; the true model adds a
    Gaussian anomaly (sigma 200 m, +10 % vp and vs) centred at a point of
    seed 6 (x in the middle half, depth in the middle 40 %).  nshots
    explosive Ricker shots evenly spaced along x at 125 m depth, fc ~ U(8, 20)
    Hz (seed 7), amplitude amp; 300 receivers at 115 m depth.  The
    gradient scales with the square of the source amplitude: the norm of the
    middle shot's gradient (3 shots, NT 2000) is 1.1e-22 at amplitude 1 and
    1.2e-4 at the default 1e9, which keeps the gradient files' values away
    from the bottom of the float range.  which = "true": MODE=FW on
    the true model writing the measured data ./data/obs_{vx,vy}.su.shot<n>;
    "init": MODE=FWI (LNORM 2, VELOCITY 1, ADJOINT_TYPE 1, DTINV) on the
    background reading them."""
    dh, dt, fw = 5.0, 5.0e-4, 20
    depth = np.arange(ny) * dh
    vp = np.repeat((2000.0 + 2000.0 * depth / depth[-1])[None, :], nx, axis=0)
    rng = np.random.Generator(np.random.PCG64(6))
    cx = rng.uniform(0.25, 0.75) * nx * dh
    cz = rng.uniform(0.3, 0.7) * ny * dh
    x = (np.arange(nx) * dh - cx)[:, None]
    z = (np.arange(ny) * dh - cz)[None, :]
    bump = 1.0 + 0.10 * np.exp(-(x * x + z * z) / (2.0 * 200.0 ** 2))
    v = vp * (bump if which == "true" else 1.0)
    model = {"vp": v.astype(np.float32), "vs": (v / 1.73).astype(np.float32),
             "rho": (310.0 * vp ** 0.25).astype(np.float32)}
    fcs = np.random.Generator(np.random.PCG64(7)).uniform(8.0, 20.0, nshots)
    x0, x1 = (fw + 10) * dh, (nx - fw - 11) * dh
    xs = np.linspace(x0, x1, nshots)
    srcs = [(float(xx) + 1.0, 125.0 + 1.0, 0.0, float(f), float(amp)) for xx, f in zip(xs, fcs)]
    recs = [(float(xx) + 1.0, 115.0 + 1.0) for xx in np.linspace((fw + 5) * dh, (nx - fw - 6) * dh, 300).round() // dh * dh]
    fwi = which == "init"
    extra = dict(FWI_KEYS, DATA_DIR="./data/obs", ADJOINT_TYPE=1, VELOCITY=1, DTINV=dtinv) if fwi else {}
    c = Case(name=f"c5_{which}", nx=nx, ny=ny, dh=dh, dt=dt, time=nt * dt, weq="EL_ISO", fdorder=8, maxrelerror=1,
             abs_type=1, fw=fw, free_surf=0, fpml=14.0, vppml=4000.0, source_type=1, sources=srcs, receivers=recs,
             seismo=1, ndt=1, run_multiple_shots=1, mode="FWI" if fwi else "FW", extra=extra)
    js = write_case(d, c, model)
    os.makedirs(os.path.join(d, "data"), exist_ok=True)
    os.makedirs(os.path.join(d, "grad"), exist_ok=True)
    if not fwi:   # the measured data are this run's seismograms
        txt = open(js).read().replace('"SEIS_FILE" : "./out/seis"', '"SEIS_FILE" : "./data/obs"')
        open(js, "w").write(txt)
    return js
