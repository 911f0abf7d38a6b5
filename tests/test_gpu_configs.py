"""The large benchmark configurations on the GPU at their full sizes.

* (H2O)_256 x 8, PM3: the mixed-precision SP2 SCF against the FP64 diagonalisation SCF of the same
  inputs, |dHf| <= 1e-4 eV/atom; the SP2 density has trace 2 nocc (to 1e-8 relative) and is
  idempotent (max |P^2 - 2P| <= 1e-5) in FP64.
* (H2O)_1024, PM3 (6144 orbitals): the SP2 SCF converges; trace and idempotency as above.
* The serving mix at 2000 requests/s (AM1 alkanes n <= 30 and PM3 water clusters of 4-256
  molecules) beside 256 XL-BOMD trajectories: every request converges and a sample of requests
  equals the same molecule computed alone (1e-5 eV/atom, the mixed density's SCF tolerance).
* An SP2 XL-BOMD trajectory (the serving loop's fixed-iteration step) against the oracle's
  diagonalisation XL-BOMD from the same start, Etot(t) within 1e-4 eV.
"""
import threading
import time

import numpy as np
import pytest
import torch

from pyseqm_amd import workloads as W
from pyseqm_amd.engine import SCFEngine
from pyseqm_amd.serving import Request, SCFServer, XLService

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _check_density(eng, P):
    """Trace 2 nocc to 1e-8 relative and max |P^2 - 2P| <= 1e-5.  SP2 stops at its first iterate
    on the FP32 rounding floor (idempotency error ~1e-6 per orbital), and the McWeeny step squares
    the remaining eigenvalue deviations."""
    b = eng.b
    for m in range(b.nmol):
        n, o = int(b.norb[m]), int(b.mol_mat_off[m])
        p = P[o:o + n * n].view(n, n)
        tr = torch.trace(p).item()
        idem = (p @ p - 2.0 * p).abs().max().item()
        print(f"block {m}: n {n} |tr P - 2 nocc| {abs(tr - 2 * b.nocc[m]):.2e} max|P^2 - 2P| {idem:.2e}", flush=True)
        assert abs(tr - 2 * b.nocc[m]) < 1e-8 * 2 * b.nocc[m], (m, tr, 2 * b.nocc[m])
        assert idem < 1e-5, (m, idem)


def test_water256x8_sp2_vs_fp64():
    sp, x = W.water_clusters(256, 8, seed=0)
    print("(H2O)_256 x 8 PM3: FP64 diagonalisation SCF", flush=True)
    e1 = SCFEngine(sp, x, "PM3", DEV)
    r1 = e1.scf(eps=1e-8, converger=(2,), density="eigh")
    h1 = e1.energies(r1.P)["Hf"].cpu().numpy()
    assert not r1.notconverged.any()
    print("(H2O)_256 x 8 PM3: SP2 SCF", flush=True)
    e2 = SCFEngine(sp, x, "PM3", DEV)
    r2 = e2.scf(eps=1e-5, converger=(2,), density="sp2")
    h2 = e2.energies(r2.P)["Hf"].cpu().numpy()
    assert not r2.notconverged.any()
    err = np.abs(h2 - h1) / e2.b.nat
    print(f"SP2 vs FP64 |dHf|/atom max {err.max():.2e} mean {err.mean():.2e} eV "
          f"(iterations {r2.iterations} vs {r1.iterations})", flush=True)
    assert err.max() <= 1e-4
    _check_density(e2, r2.P)


def test_water1024_sp2_converges():
    sp, x = W.water_clusters(1024, 1, seed=0)
    print("(H2O)_1024 PM3: SP2 SCF", flush=True)
    eng = SCFEngine(sp, x, "PM3", DEV)
    r = eng.scf(eps=1e-5, converger=(2,), density="sp2")
    assert not r.notconverged.any()
    hf = eng.energies(r.P)["Hf"].cpu().numpy()
    assert np.isfinite(hf).all()
    _check_density(eng, r.P)


def test_serving_mix_at_2000_per_second():
    stream = W.serving_stream(rate=2000.0, window=0.25, seed=11)
    assert len(stream) >= 400
    sp, c = W.alkanes(256, seed=7)
    print(f"serving: {len(stream)} requests over 0.25 s beside 256 XL-BOMD trajectories", flush=True)
    xl = XLService(sp, c, "AM1", DEV, k=6)
    srv = SCFServer(DEV, lanes=6, base_lanes=3, density="sp2", eps=1e-5, max_batch=256)
    stop = threading.Event()
    th = threading.Thread(target=xl.run_until, args=(stop,), daemon=True)
    th.start()
    reqs = []
    t0 = time.perf_counter()
    for k, (ta, s, x, method) in enumerate(stream):
        while time.perf_counter() - t0 < ta:
            time.sleep(1e-4)
        reqs.append(srv.submit(Request(k, s, x, method)))
    srv.close()
    stop.set()
    th.join()
    assert xl.steps > 2
    assert srv.batches >= 4
    for r in reqs:
        assert r.result is not None and r.result["converged"], r.rid
    rng = np.random.default_rng(0)
    for i in rng.choice(len(reqs), 12, replace=False):
        r = reqs[i]
        e = SCFEngine(r.species[None], r.coords[None], r.method, DEV)
        res = e.scf(eps=1e-5, converger=(2,), density="sp2")
        hf = float(e.energies(res.P)["Hf"][0])
        nat = int((r.species > 0).sum())
        assert abs(hf - r.result["Hf"]) < 1e-5 * nat, (r.rid, r.method, hf, r.result)


def test_sp2_xl_step_matches_oracle_trajectory():
    """The serving loop's XL-BOMD step (SP2 density at a fixed iteration count, graph replays) on one
    alkane against the oracle's XL-BOMD (FP64 diagonalisation) from the same coordinates and
    velocities; the service's constructor takes the first step."""
    from pyseqm_amd.seqm_functions.constants import MASS
    from pyseqm_amd.xlbomd import VEL_SCALE
    from tests.test_gpu_xlbomd import _oracle_xl
    sp, c = W.alkanes(1, seed=4, nmin=5, nmax=5)
    xl = XLService(sp, c, "AM1", DEV, k=6, graphs=True, seed=0)
    b = xl.traj.eng.b
    rng = np.random.default_rng(0)  # the service's velocity draw (seed 0, 300 K)
    v = rng.normal(size=(b.natoms, 3)) * np.sqrt(300.0 / np.take(MASS, b.Z))[:, None] * VEL_SCALE
    v0 = b.atom_to_padded(torch.as_tensor(v, device=DEV)).cpu().numpy()
    nsteps = 6
    ours = [xl.energies()[0]]
    for _ in range(nsteps - 1):
        xl.step()
        ours.append(xl.energies()[0])
    ref = _oracle_xl(sp, c, v0, "AM1", 6, nsteps, eps=1e-8)
    print("SP2 XL-BOMD vs oracle Etot(t):", np.array(ours) - ref[1:, 1], flush=True)
    np.testing.assert_allclose(ours, ref[1:, 1], atol=1e-4, rtol=0)
