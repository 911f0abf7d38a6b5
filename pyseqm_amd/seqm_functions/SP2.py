"""SP2(a, nocc, eps, factor): trace-correcting purification of dense (N, D, D) Fock matrices
(seqm/seqm_functions/SP2.py:9-85) on the mixed-precision MFMA kernels.

Every D x D block is one system, zero rows included, exactly as the reference treats its input.
X runs in FP32 with X^2 on bf16x3 MFMA (FP32 accumulation); a block stops when its trace error is
below max(eps, 1e-2) (SP2.py:21-28) and its idempotency error e = tr X - tr X^2 no longer falls
quadratically (e_k >= 8 e_{k-2}^2 once e_{k-2} < 3/16), in place of the reference's 'trace error
no longer improving' (SP2.py:74-78), which the cancelling eigenvalue deviations can fool; the returned
density gets one McWeeny step to FP64 accuracy, factor * (3 X^2 - 2 X^3)."""
import numpy as np
import torch

from .. import _native as nat
from ..layout import MatrixLayout
from ..sp2 import SP2_EPS, density_sp2


class _MatrixEngine:
    def __init__(self, b):
        self.b = b
        self.device = b.device
        self.stream = nat.stream_ptr(b.device)
        self.red = torch.empty(b.nmol * nat.MAX_RED_BLOCKS * 16, dtype=torch.float64, device=b.device)
        self._idx_host = np.arange(b.nmol)


def SP2(a, nocc, eps=1.0e-4, factor=2.0):
    dev = nat.require_device(a.device)
    N, D, _ = a.shape
    b = MatrixLayout(N, D, nocc, dev)
    eng = _MatrixEngine(b)
    F = a.detach().to(torch.float64).contiguous().reshape(-1)
    P = torch.empty_like(F)
    idx = torch.arange(N, dtype=torch.int32, device=dev)
    density_sp2(eng, F, idx, N, P, eps=max(float(eps), SP2_EPS))
    return (0.5 * factor * P).reshape(N, D, D).to(a.dtype)
