"""Mixed-precision SP2 density step on MFMA (host driver of the sp2.hip kernels).

Replaces pack -> SP2 -> unpack of make_Pnew_factory (seqm/seqm_functions/scf_loop.py:66-78,
SP2.py:9-85): trace-correcting SP2 with X in FP32 and X^2 on bf16x3 MFMA (FP32 accumulate),
branch / stop decided on the device, then one McWeeny step P = 2 (3 X^2 - 2 X^3) to FP64
accuracy: X^2 on the FP64 MFMA GEMM, X^3 = X^2 + X (X^2 - X) with the small correction product on
bf16x3 MFMA.  Stop rule: trace error below eps (SP2.py:75-79) and the idempotency error
e = tr X - ||X||_F^2 no longer falling quadratically (e_k >= 8 e_{k-2}^2 once e_{k-2} < 3/16: the
first iterate on the FP32 rounding floor); the reference's own reduced-precision rule watches the
trace error, in which occupied and virtual deviations cancel, and can stop mid-run with
eigenvalues still ~1e-2 from 0 or 1.
"""
import numpy as np
import torch

from . import _native as nat

NSTATE = 8
SP2_EPS = 1.0e-2      # SP2.py:4 (reduced-precision threshold)
SP2_MAX_ITER = 100
POLL = 4              # iterations between host checks of the stop counter


class SP2Workspace:
    def __init__(self, eng):
        b = eng.b
        dev = eng.device
        ld = ((b.norb + 31) // 32 * 32).astype(np.int64)
        off = np.concatenate([[0], np.cumsum(ld * ld)]).astype(np.int64)
        self.xtotal = int(off[-1])
        self.xoff = torch.as_tensor(off[:-1], device=dev)
        self.xld = torch.as_tensor(ld.astype(np.int32), device=dev)
        self.Xa = torch.zeros(max(self.xtotal, 1), dtype=torch.float32, device=dev)
        self.Xb = torch.zeros_like(self.Xa)
        self.Y2 = torch.zeros(max(self.xtotal, 1), dtype=torch.float64, device=dev)  # McWeeny X^2
        self.state = torch.zeros(b.nmol, NSTATE, dtype=torch.float64, device=dev)
        # per-tile (trace, Frobenius) partials; sized for the 64 x 64 tiles (the 128 x 128 tiles
        # used for large blocks need fewer)
        nt = (b.nmax + 63) // 64
        self.part_stride = max(nt * (nt + 1), 2)
        self.part = torch.zeros(b.nmol, self.part_stride, dtype=torch.float64, device=dev)
        self.ndone = torch.zeros(1, dtype=torch.int32, device=dev)
        # two pinned slots: the stop counter of chunk c is copied while chunk c+1 is launched
        self.ndone_host = torch.zeros(2, dtype=torch.int32, pin_memory=True)
        self.last_iters = None
        self.timing = None  # list of (start event, end event, idx, nact) of the SP2 GEMM loops


def _xcd_lpt_order(eng):
    """The active molecules by decreasing size, dealt round-robin into 8 runs (the SP2 kernels give
    each XCD a contiguous run of the molecule-major tile list): every XCD gets an even share, and
    within its run the longest K loops start first, so the short ones fill the tail.  Only the
    launch order changes; every kernel indexes its data by molecule."""
    ih = np.asarray(eng._idx_host)
    srt = ih[np.argsort(-eng.b.norb[ih], kind="stable")]
    perm = np.concatenate([srt[x::8] for x in range(8)]).astype(np.int32)
    return torch.as_tensor(perm, device=eng.device)


def density_sp2(eng, F, idx, nact, Pout, eps=SP2_EPS, max_iter=SP2_MAX_ITER, fixed_iters=None):
    """fixed_iters: launch exactly that many SP2 iterations and never read the stop counter on the
    host (molecules that stopped make the remaining launches no-ops); the launch sequence is then
    free of host synchronisation and can be captured in a graph."""
    b = eng.b
    ws = getattr(eng, "_sp2ws", None)
    if ws is None:
        ws = eng._sp2ws = SP2Workspace(eng)
    nmax = int(b.norb[eng._idx_host].max())
    if fixed_iters is None and nact > 8 and not torch.cuda.is_current_stream_capturing():
        idx = _xcd_lpt_order(eng)
    s = eng.stream
    L = b.L
    nat.call("pyseqm_sp2_init", L, nat.ptr(F), nat.ptr(idx), nact, nat.ptr(ws.xoff), nat.ptr(ws.xld),
             nat.ptr(ws.Xa), nat.ptr(ws.state), nat.ptr(eng.red), nat.red_blocks(nmax), s)
    ws.ndone.zero_()
    if ws.timing is not None:
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record()
    args = (nat.ptr(idx), nact, nmax, nat.ptr(ws.xoff), nat.ptr(ws.xld), nat.ptr(ws.Xa), nat.ptr(ws.Xb),
            nat.ptr(ws.state), nat.ptr(ws.part), ws.part_stride, float(eps), int(max_iter), nat.ptr(ws.ndone))

    def launch_many(it0, count):
        nat.call("pyseqm_sp2_iterate_many", L, *args, it0, count, s)

    if fixed_iters is not None:
        launch_many(0, int(fixed_iters))
    else:
        # pipelined stop test: after each chunk of POLL iterations the counter is copied to pinned
        # host memory behind an event; the host reads the previous chunk's copy while the current
        # chunk runs, so the stream never drains for the test (molecules that stopped make the
        # extra chunk's launches no-ops)
        it, chunk, pending = 0, 0, None
        while it < max_iter:
            k = min(POLL, max_iter - it)
            launch_many(it, k)
            it += k
            slot = chunk % 2
            ws.ndone_host[slot:slot + 1].copy_(ws.ndone, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record()
            if pending is not None:
                pev, pslot = pending
                pev.synchronize()
                if int(ws.ndone_host[pslot]) >= nact:
                    break
            pending = (ev, slot)
            chunk += 1
    if ws.timing is not None:
        ev1.record()
        ws.timing.append((ev0, ev1, idx.clone(), nact))
    # McWeeny correction P = 2 (3 X^2 - 2 X^3), as P = 2 (Y2 - 2 X D) with Y2 = X^2 in FP64 and
    # D = Y2 - X (small) multiplied on bf16x3: FP64-accurate P
    nat.call("pyseqm_sp2_mcweeny", L, nat.ptr(idx), nact, nmax, nat.ptr(ws.xoff), nat.ptr(ws.xld), nat.ptr(ws.state),
             nat.ptr(ws.Xa), nat.ptr(ws.Xb), nat.ptr(ws.Y2), nat.ptr(Pout), s)
    ws.last_iters = ws.state[:, 7]
    if ws.timing is not None:
        ws.timing[-1] = ws.timing[-1] + (ws.state[:, 7].clone(),)


def gemm_flops(eng, idx, iters):
    """Algorithmic FLOPs of the SP2 X^2 products: 2 n^3 per iteration per molecule."""
    n = torch.as_tensor(eng.b.norb, device=iters.device, dtype=torch.float64)
    il = idx.long()
    return float((2.0 * n[il] ** 3 * iters[il]).sum())


SMALL_TILE_MAX_N = 256  # largest active block for which the SP2 kernels use 64 x 64 output tiles


def executed_mfma_flops(eng, idx, iters):
    """FLOPs the SP2 kernels issue on the MFMA pipe: per molecule and iteration, the upper-triangle
    output tiles (edge TB = 64 while the largest active block has <= 256 orbitals, else 128) that
    hold rows of the block, each a K loop over the padded leading dimension, three bf16 products."""
    il = idx.long().cpu().numpy()
    n = eng.b.norb[il].astype(np.float64)
    tb = 64 if n.max() <= SMALL_TILE_MAX_N else 128
    nt = np.ceil(n / tb)
    ld = np.ceil(n / 32) * 32
    it = iters.double().cpu().numpy()[il]
    return float((3.0 * 2.0 * tb * tb * ld * nt * (nt + 1) / 2 * it).sum())
