"""two_elec_two_center_int: molecular-frame two-center integrals w, electron-core attractions
e1b / e2a and rho0 of an explicit pair list (seqm/seqm_functions/two_elec_two_center_int.py:98-283,
e1b/e2a layout 1534-1569), on the pair-integral kernel.

Each pair becomes a two-atom system (atom i at the origin, atom j at x_ij r_ij), so the result
depends only on the pair data passed in, like the reference; per-atom parameters come from the
arguments (learned values included).  riXH / ri are returned as None."""
import numpy as np
import torch

from .. import _native as nat
from .constants import TORE, a0, ev

_TRI_I = (0, 0, 1, 0, 1, 2, 0, 1, 2, 3)
_TRI_J = (0, 1, 1, 2, 2, 2, 3, 3, 3, 3)


def two_elec_two_center_int(const, idxi, idxj, ni, nj, xij, rij, Z, zetas, zetap, zetad, zs, zp, zd, gss, gpp, gp2,
                            hsp, F0SD, G2SD, rho_core, alpha, chi, themethod):
    from ..engine import SCFEngine
    dev = nat.require_device(xij.device)
    npairs = int(xij.shape[0])
    ii, jj = idxi.long(), idxj.long()
    sp = torch.stack((ni.long(), nj.long()), 1)
    xyz = torch.zeros(npairs, 2, 3, dtype=torch.float64, device=dev)
    xyz[:, 1] = xij.to(torch.float64) * (rij.to(torch.float64) * a0).unsqueeze(1)
    from ..layout import PackedBatch
    tore_np = np.asarray(TORE)
    spn = sp.cpu().numpy()
    # two-atom systems for the integrals only: any electron count is fine
    b = PackedBatch(spn, themethod, dev, charges=tore_np[spn].sum(1).astype(np.int64) % 2)
    eng = SCFEngine(spn, None, themethod, dev, batch=b)
    over = {"zeta_s": zetas, "zeta_p": zetap, "g_ss": gss, "g_pp": gpp, "g_p2": gp2, "h_sp": hsp}
    eng.set_parameters({k: torch.stack((v[ii], v[jj]), 1).reshape(-1) for k, v in over.items() if v is not None})
    eng.set_coordinates(xyz)
    w = eng.w_reference()
    tore = torch.as_tensor(np.asarray(TORE), dtype=torch.float64, device=dev)
    ti, tj = tore[ni.long()], tore[nj.long()]
    e1b = torch.zeros(npairs, 4, 4, dtype=torch.float64, device=dev)
    e2a = torch.zeros_like(e1b)
    for k in range(10):
        e1b[:, _TRI_I[k], _TRI_J[k]] = -tj * w[:, k, 0]
        e2a[:, _TRI_I[k], _TRI_J[k]] = -ti * w[:, 0, k]
    rho0 = 0.5 * ev / gss.to(torch.float64)
    return w, e1b, e2a, rho0[ii], rho0[jj], None, None
