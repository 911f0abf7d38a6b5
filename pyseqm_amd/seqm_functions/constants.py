"""Physical constants and element tables (seqm/seqm_functions/constants.py:4-223).

Tables cover Z = 0..18 (H..Ar), the elements an sp-only MNDO/AM1/PM3 calculation can use.
"""
import torch

ev = 27.21            # eV per hartree used by MOPAC7 (constants.py:4)
a0 = 0.529167         # bohr radius, Angstrom (constants.py:9)
ev_kcalpmol = 23.061  # kcal/mol per eV (constants.py:11)
overlap_cutoff = 40.0  # bohr (constants.py:23)

# valence core charge, principal quantum number, isolated-atom energy coefficients,
# experimental heats of formation (kcal/mol) and masses, per Z
TORE = [0, 1, 0, 1, 2, 3, 4, 5, 6, 7, 0, 1, 2, 3, 4, 5, 6, 7, 0]
QN = [0, 1, 1, 2, 2, 2, 2, 2, 2, 2, 2, 3, 3, 3, 3, 3, 3, 3, 3]
USSC = [0, 1, 0, 1, 2, 2, 2, 2, 2, 2, 0, 1, 2, 2, 2, 2, 2, 2, 0]
UPPC = [0, 0, 0, 0, 0, 1, 2, 3, 4, 5, 6, 0, 0, 1, 2, 3, 4, 5, 6]
GSSC = [0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 0, 0, 1, 1, 1, 1, 1, 1, 0]
GSPC = [0, 0, 0, 0, 0, 2, 4, 6, 8, 10, 0, 0, 0, 2, 4, 6, 8, 10, 0]
HSPC = [0, 0, 0, 0, 0, -1, -2, -3, -4, -5, 0, 0, 0, -1, -2, -3, -4, -5, 0]
GP2C = [0, 0, 0, 0, 0, 0, 1.5, 4.5, 6.5, 10, 0, 0, 0, 0, 1.5, 4.5, 6.5, 10, 0]
GPPC = [0, 0, 0, 0, 0, 0, -0.5, -1.5, -0.5, 0, 0, 0, 0, 0, -0.5, -1.5, -0.5, 0, 0]
EHEAT = [0.0, 52.102, 0.0, 38.410, 76.960, 135.700, 170.890, 113.000, 59.559, 18.890, 0.0,
         25.850, 35.000, 79.490, 108.390, 75.570, 66.400, 28.990, 0.0]
MASS = [0.0, 1.00790, 4.00260, 6.94000, 9.01218, 10.81000, 12.01100, 14.00670, 15.99940, 18.99840,
        20.17900, 22.98977, 24.30500, 26.98154, 28.08550, 30.97376, 32.06000, 35.45300, 39.94800]
MAX_Z = len(TORE) - 1


class Constants(torch.nn.Module):
    """Element tables as buffers, drop-in for seqm.seqm_functions.constants.Constants."""

    def __init__(self, do_timing=False, length_conversion_factor=(1.0 / a0), energy_conversion_factor=1.0):
        super().__init__()
        self.length_conversion_factor = length_conversion_factor
        self.energy_conversion_factor = energy_conversion_factor
        f = torch.get_default_dtype()
        for name, tab in (("tore", TORE), ("qn", QN), ("ussc", USSC), ("uppc", UPPC), ("gssc", GSSC),
                          ("gspc", GSPC), ("hspc", HSPC), ("gp2c", GP2C), ("gppc", GPPC), ("mass", MASS)):
            self.register_buffer(name, torch.tensor(tab, dtype=f))
        self.register_buffer("eheat", torch.tensor(EHEAT, dtype=f) / ev_kcalpmol)
        self.register_buffer("qn_int", torch.tensor(QN, dtype=torch.int64))
        self.do_timing = do_timing
        if do_timing:
            self.timing = {"Hcore + STO Integrals": [], "SCF": [], "Force": [], "MD": [], "D*": []}

    def forward(self):
        return None
