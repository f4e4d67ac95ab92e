"""The reference-side binding: the reference built from its own sources with
its time_loop() (src/time_loop.c:27-326) replaced by time_loop_mi355x()
(oracle/ref/time_loop_mi355x.c), which drives libsofi_hip.so through
include/sofi_hip.h.  The reference's own parameter reading, model setup,
source and receiver setup and output writers run around the GPU loop; its
seismograms and snapshot files are held to the one-process reference runs
of tests/golden/ref1 at the parity tolerance."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import cases
import common
from common import COMPS, TOL

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "sofi2D_ref_hip")
MPIEXEC = "/opt/conda/bin/mpiexec"

NAMES = (list(cases.CASES) + list(cases.FDT4_CASES) + list(cases.AC_CASES) + list(cases.TOPO_CASES) +
         list(cases.PERIODIC_CASES) + list(cases.BUILTIN_CASES))


def _read_su(path):
    b = open(path, "rb").read()
    ns = int(np.frombuffer(b[114:116], np.uint16)[0])
    tr = 240 + 4 * ns
    return np.stack([np.frombuffer(b[k * tr + 240:(k + 1) * tr], np.float32) for k in range(len(b) // tr)])


@pytest.mark.skipif(not (os.path.exists(BIN) and os.path.exists(MPIEXEC)),
                    reason="the reference build with the binding (oracle/_ref/sofi2D_ref_hip) or MPICH is absent")
@pytest.mark.parametrize("name", NAMES)
def test_reference_with_gpu_time_loop(name):
    td = tempfile.mkdtemp(prefix="sofi_bind_")
    js, c = cases.build(name, td)
    print("binding case", name, flush=True)
    r = subprocess.run(["timeout", "-k", "10", "300", MPIEXEC, "-np", "1", BIN, js], cwd=td, capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    ref = common.ref1(name)
    errs = {}
    if c.run_multiple_shots:
        for s in range(1, len(c.sources) + 1):
            comps = {1: ("vx", "vy"), 4: COMPS}[c.seismo]
            pks = [ref[f"shot{s}_seis_{k}"] for k in comps if k in ("vx", "vy", "p")]
            for comp in comps:
                a = _read_su(os.path.join(td, "out", f"seis_{comp}.su.shot{s}"))
                e = common.compare_component(a, ref[f"shot{s}_seis_{comp}"], pks, TOL, None)
                errs[f"shot{s}_{comp}"] = e
                assert e is None or e <= TOL, (s, comp, e)
    else:
        iso_expl = c.weq in ("EL_ISO", "VEL_ISO") and c.source_type == 1
        secs = {k[5:]: v for k, v in ref.items() if k.startswith("seis_")}
        snaps = {k[5:]: v for k, v in ref.items() if k.startswith("snap_")}
        pks = [secs[k] for k in ("vx", "vy", "p") if k in secs]
        for comp, g in secs.items():
            a = _read_su(os.path.join(td, "out", f"seis_{comp}.su"))
            cancel = float(np.abs(secs["div"]).max()) if (iso_expl and comp == "curl" and "div" in secs) else None
            e = common.compare_component(a, g, pks, TOL, cancel)
            errs["seis_" + comp] = e
            assert e is None or e <= TOL, (comp, e)
        spk = [snaps[k] for k in ("vx", "vy", "p") if k in snaps]
        for comp, g in snaps.items():
            a = np.fromfile(os.path.join(td, "out", f"snap.bin.{comp}.0.0"), np.float32).reshape(g.shape)
            cancel = float(np.abs(snaps["div"]).max()) if (iso_expl and comp == "curl" and "div" in snaps) else None
            e = common.compare_component(a, g, spk, TOL, cancel)
            errs["snap_" + comp] = e
            assert e is None or e <= TOL, (comp, e)
    print(name, errs)
