"""Single-GPU serving loop: continuous batching of SCF single-point requests plus a persistent,
graph-captured batch of XL-BOMD trajectories, all resident in HBM.

* SCF requests (species, coordinates, method, charge) enter a FIFO queue with their arrival time.
  ``lanes`` worker threads, each owning one HIP stream (default priority), repeatedly take the
  oldest admissible group of same-method requests, pack it padding-free (PackedBatch), and run the
  device SCF engine to convergence.  A request is admissible while the HBM footprint of all
  batches in flight plus its own stays under the budget; requests that arrive while other
  batches run are admitted as soon as a lane and memory free up, and their kernels overlap with
  the running batches on the other streams.  Each result is copied to the host and the request
  is completed with its arrival->result latency.
* XL-BOMD trajectories (one step per request, closed loop: every trajectory issues its next step
  when the previous one returns) live in one ``XLTrajectories`` batch: coordinates, velocities,
  density P, the P history ring and D stay in HBM between steps.  The step contains no host
  synchronisation (fixed SP2 iteration count), so it is captured once per ring position
  (k + 1 graphs) and replayed.

The reference has no serving layer; the per-request computation is its single point
(Force(..., do_force=False), basics.py:1247-1365) and XL_BOMD.one_step
(MolecularDynamics.py:1437-1515).
"""
import collections
import threading
import time

import numpy as np
import torch

from .engine import SCFEngine
from .layout import PackedBatch


def footprint_bytes(norb, npairs, density="sp2"):
    """HBM bytes one molecule of norb orbitals / npairs atom pairs holds during its SCF: ~30 FP64
    n^2 matrices (P, F, H, mixing copies, 10-deep Pulay ring of Fock matrices and residuals), the
    compact pair integrals (<= 100 doubles per pair) and the density-step workspace (SP2: FP32
    X ping-pong + 3 FP64 padded blocks; eigh: padded input, vectors and solver workspace)."""
    n2 = float(norb) ** 2
    ld = (norb + 31) // 32 * 32
    work = 32.0 * ld * ld if density == "sp2" else 40.0 * n2
    return int(8 * 30 * n2 + 8 * 110 * npairs + work)


class Request:
    __slots__ = ("rid", "species", "coords", "method", "charge", "t_arrival", "t_admit", "t_done", "result",
                 "event", "norb", "npairs", "bytes")

    def __init__(self, rid, species, coords, method, charge=0, t_arrival=None):
        sp = np.asarray(species, dtype=np.int64).reshape(-1)
        order = np.argsort(-sp, kind="stable")  # the reference's species order: non-increasing Z
        self.rid = rid
        self.species = sp[order]
        self.coords = np.asarray(coords, dtype=np.float64).reshape(-1, 3)[order]
        self.method = method
        self.charge = int(charge)
        self.t_arrival = time.perf_counter() if t_arrival is None else t_arrival
        self.t_admit = self.t_done = None
        self.result = None
        self.event = threading.Event()
        nat = int((self.species > 0).sum())
        self.norb = int(4 * (self.species > 1).sum() + (self.species == 1).sum())
        self.npairs = nat * (nat - 1) // 2
        self.bytes = 0

    @classmethod
    def from_records(cls, buf, first_rid=0, t_arrival=None):
        """Requests from a stream of binary request records (pyseqm_amd.wire)."""
        from . import wire
        return [cls(first_rid + k, Z, X, method, charge, t_arrival)
                for k, (Z, X, method, charge) in enumerate(wire.decode(buf))]

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.t_arrival


class SCFServer:
    """Continuous-batching SCF service on one GPU (see module docstring)."""

    def __init__(self, device="cuda:0", lanes=3, density="sp2", eps=1.0e-5, converger=(2,),
                 hbm_budget=None, batch_bytes=None, max_batch=512, base_lanes=None, backlog=256):
        self.device = torch.device(device)
        self.density = density
        self.eps = float(eps)
        self.converger = tuple(converger)
        from . import _native
        _native.lib()
        free, total = torch.cuda.mem_get_info(self.device)
        self.hbm_budget = int(hbm_budget if hbm_budget is not None else 0.6 * free)
        self.batch_bytes = int(batch_bytes if batch_bytes is not None else min(8 << 30, self.hbm_budget // lanes))
        self.max_batch = int(max_batch)
        # lanes beyond base_lanes only take work while more than `backlog` requests wait: under
        # light load a few larger batches keep latency low, under overload the extra concurrent
        # batches overlap their host work and kernels and raise throughput
        self.base_lanes = int(lanes if base_lanes is None else min(base_lanes, lanes))
        self.backlog = int(backlog)
        self.queue = collections.deque()
        self.cv = threading.Condition()
        self.in_use = 0
        self.stop = False
        self.completed = []
        self.batches = 0
        self.peak_in_use = 0
        self.errors = []
        self.threads = [threading.Thread(target=self._lane, args=(k,), daemon=True) for k in range(lanes)]
        for t in self.threads:
            t.start()

    # ----------------------------------------------------------------- API
    def submit(self, req):
        req.bytes = footprint_bytes(req.norb, req.npairs, self.density)
        if req.bytes > self.hbm_budget:
            raise ValueError(f"request {req.rid}: footprint {req.bytes / 2**30:.1f} GiB exceeds the HBM budget")
        with self.cv:
            self.queue.append(req)
            self.cv.notify()
        return req

    def close(self):
        with self.cv:
            self.stop = True
            self.cv.notify_all()
        for t in self.threads:
            t.join()
        if self.errors:
            raise self.errors[0]

    # ----------------------------------------------------------- admission
    @staticmethod
    def size_class(norb):
        """Blocks up to 256 orbitals run on the small-block kernels (64 x 64 SP2 tiles, FP64 DIIS
        residual); a batch that mixes them with larger blocks would run every block on the
        large-block ones, so batches keep to one class."""
        return 0 if norb <= 256 else 1

    def _take(self):
        """Oldest request first; then every queued request of the same method and size class that
        fits the batch cap, as long as the in-flight footprint stays under the HBM budget."""
        q = self.queue
        head = q[0]
        if self.in_use + head.bytes > self.hbm_budget:
            return None
        cls = self.size_class(head.norb)
        batch, total, keep = [head], head.bytes, collections.deque()
        q.popleft()
        while q and len(batch) < self.max_batch:
            r = q.popleft()
            if (r.method == head.method and self.size_class(r.norb) == cls and total + r.bytes <= self.batch_bytes
                    and self.in_use + total + r.bytes <= self.hbm_budget):
                batch.append(r)
                total += r.bytes
            else:
                keep.append(r)
        keep.extend(q)
        self.queue = keep
        self.in_use += total
        self.peak_in_use = max(self.peak_in_use, self.in_use)
        return batch, total

    def _lane(self, k):
        stream = torch.cuda.Stream(device=self.device)
        while True:
            with self.cv:
                got = None
                while not self.stop or self.queue:
                    if self.queue and (k < self.base_lanes or len(self.queue) > self.backlog or self.stop):
                        got = self._take()
                        if got is not None:
                            break
                    self.cv.wait(timeout=0.01)
                if got is None:
                    return
            batch, nbytes = got
            try:
                with torch.cuda.stream(stream):
                    self._run(batch)
            except Exception as e:  # surfaced by close()
                self.errors.append(e)
                for r in batch:
                    r.event.set()
            finally:
                with self.cv:
                    self.in_use -= nbytes
                    self.batches += 1
                    self.cv.notify_all()

    def _run(self, batch):
        t_admit = time.perf_counter()
        K = max(len(r.species) for r in batch)
        sp = np.zeros((len(batch), K), dtype=np.int64)
        xyz = np.zeros((len(batch), K, 3))
        for i, r in enumerate(batch):
            sp[i, :len(r.species)] = r.species
            xyz[i, :len(r.species)] = r.coords
            r.t_admit = t_admit
        charges = np.array([r.charge for r in batch])
        b = PackedBatch(sp, batch[0].method, self.device, charges if charges.any() else None)
        eng = SCFEngine(sp, None, batch[0].method, self.device, batch=b)
        eng.set_coordinates(torch.as_tensor(xyz, device=self.device))
        res = eng.scf(eps=self.eps, converger=self.converger, density=self.density)
        en = eng.energies(res.P)
        host = {k: en[k].cpu().numpy() for k in ("Hf", "Etot", "Eelec", "Enuc")}
        t_done = time.perf_counter()
        for i, r in enumerate(batch):
            r.result = {k: float(v[i]) for k, v in host.items()}
            r.result["converged"] = not bool(res.notconverged[i])
            r.result["batch_size"] = len(batch)
            r.t_done = t_done
            r.event.set()
        with self.cv:
            self.completed.extend(batch)


class XLService:
    """Persistent XL-BOMD trajectory batch stepped by graph replays (see module docstring)."""

    def __init__(self, species, coords, method, device="cuda:0", k=6, dt=0.5, temp=300.0, seed=0,
                 sp2_iters=48, scf_eps=1.0e-6, graphs=True):
        from .xlbomd import VEL_SCALE, XLTrajectories
        self.device = torch.device(device)
        self.stream = torch.cuda.Stream(device=self.device)
        with torch.cuda.stream(self.stream):
            eng = SCFEngine(species, coords, method, self.device)
            r = eng.scf(eps=scf_eps, converger=(2,), density="sp2")
            b = eng.b
            rng = np.random.default_rng(seed)
            from .seqm_functions.constants import MASS
            m = np.take(MASS, b.Z)
            v = rng.normal(size=(b.natoms, 3)) * np.sqrt(temp / m)[:, None] * VEL_SCALE
            x0 = b.atoms_from_padded(torch.as_tensor(coords, device=self.device))
            self.traj = XLTrajectories(eng, k, dt, x0, v, r.P, density="sp2")
            self.traj.set_forces(eng.forces(r.P))
            self.sp2_iters = int(sp2_iters)
            # warm-up eager step (allocates every workspace outside capture)
            self._body(0)
            self.graphs = []
            for cindx in range(self.traj.m if graphs else 0):
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g, stream=self.stream):
                    self._body(cindx)
                self.graphs.append(g)
        torch.cuda.synchronize(self.device)
        self.steps = 1
        self.latencies = []

    def _body(self, cindx):
        tr = self.traj
        tr.step_count = cindx
        tr._verlet(0)
        tr._propagate()
        eng, b = tr.eng, tr.eng.b
        eng.set_coordinates(tr.x)
        eng.fock(tr.P, tr.F)
        eng.density_kind = "sp2"
        eng._idx_host = np.arange(b.nmol)
        eng.density(tr.F, tr.idx_all, b.nmol, tr.D, fixed_iters=self.sp2_iters)
        from . import _native as nat
        nat.call("pyseqm_xl_energy", b.L, nat.ptr(tr.D), nat.ptr(tr.P), nat.ptr(tr.F), nat.ptr(eng.H),
                 nat.ptr(tr.eelec), nat.ptr(eng.red), nat.red_blocks(b.nmax), eng.stream)
        nat.call("pyseqm_pair_forces_xl", b.L, nat.METHODS[eng.method], nat.ptr(tr.x), nat.ptr(eng.atom_par),
                 nat.ptr(tr.D), nat.ptr(tr.P), tr.force_h, nat.ptr(eng.part), nat.ptr(tr.f), eng.stream)
        tr._verlet(1)

    def step(self):
        """Replay one step of every trajectory; returns the wall time of the step (s)."""
        t0 = time.perf_counter()
        with torch.cuda.stream(self.stream):
            if self.graphs:
                self.graphs[self.steps % self.traj.m].replay()
            else:
                self._body(self.steps % self.traj.m)
        self.stream.synchronize()
        self.steps += 1
        dt = time.perf_counter() - t0
        self.latencies.append(dt)
        return dt

    def checkpoint(self, path):
        """Write every trajectory's device state and the step counter (safetensors)."""
        from .xlbomd import save_state
        self.stream.synchronize()
        save_state(path, self.traj.state_dict(), {"service_steps": self.steps})

    def restore(self, path):
        """Load a checkpoint into the resident batch in place: the captured graphs keep replaying on
        the same buffers, and the trajectories continue from the saved step."""
        from .xlbomd import load_state
        state, extra = load_state(path)
        with torch.cuda.stream(self.stream):
            self.traj.load_state_dict(state)
        self.stream.synchronize()
        self.steps = int(extra["service_steps"])

    def run_until(self, stop_event):
        while not stop_event.is_set():
            self.step()

    def energies(self):
        with torch.cuda.stream(self.stream):
            e = self.traj.energies()
            ek = self.traj.kinetic_energy()
        self.stream.synchronize()
        return (e["Etot"] + ek).cpu().numpy()


def percentiles(x, qs=(50, 99)):
    x = np.asarray(x, dtype=np.float64)
    return {f"p{q}": float(np.percentile(x, q)) if len(x) else None for q in qs}
