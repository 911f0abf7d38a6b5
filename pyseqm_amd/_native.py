"""ctypes binding of libpyseqm_hip.so (the C ABI declared in include/pyseqm_hip.h).

The library is loaded from this package directory only.  There is no fallback: if it is
missing or the device is not a HIP GPU, every compute entry point raises.
"""
import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpyseqm_hip.so")

c_int, c_double, c_void_p, c_longlong = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_longlong

NPAR = 32
METHODS = {"MNDO": 0, "AM1": 1, "PM3": 2}


class Layout(ctypes.Structure):
    _fields_ = [
        ("nmol", c_int), ("natoms", c_int), ("ntile", c_int * 3), ("npairs", c_int),
        ("mol_atom_off", c_void_p), ("mol_norb", c_void_p), ("mol_nocc", c_void_p),
        ("mol_mat_off", c_void_p), ("mol_tile_off", c_void_p), ("atom_mol", c_void_p),
        ("atom_ao", c_void_p), ("atom_Z", c_void_p), ("tiles", c_void_p), ("csr_off", c_void_p),
        ("csr_idx", c_void_p), ("tile_woff", c_void_p),
        ("max_mol_atoms", c_int), ("max_mol_slots", c_int), ("mol_order", c_void_p),
    ]


_P = ctypes.POINTER(Layout)
# name -> argtypes (after the layout pointer); every function returns int status
_SIGS = {
    "pyseqm_atom_setup": [c_void_p, c_void_p],
    "pyseqm_pair_integrals": [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                              c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    "pyseqm_pair_forces": [c_int, c_void_p, c_void_p, c_void_p, c_double, c_void_p, c_void_p, c_void_p],
    "pyseqm_param_grad": [c_int, c_void_p, c_void_p, c_void_p, c_double, c_void_p, c_void_p, c_void_p, c_void_p],
    "pyseqm_fock_param_vjp": [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_void_p, c_void_p, c_void_p,
                              c_void_p],
    "pyseqm_density_response_elem": [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p],
    "pyseqm_fock": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                    c_void_p, c_void_p, c_void_p, c_void_p],
    "pyseqm_fock_noalloc": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                            c_void_p, c_void_p],
    "pyseqm_pad_for_eigh": [c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_int, c_void_p],
    "pyseqm_eigh_density": [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    "pyseqm_eigh_vectors": [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                            c_void_p],
    "pyseqm_fermi_density": [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_double, c_double, c_void_p,
                             c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    "pyseqm_block_gemm": [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p],
    "pyseqm_canon_prt": [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_double, c_double, c_int, c_void_p,
                         c_void_p],
    "pyseqm_block_dot": [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p],
    "pyseqm_block_axpy": [c_void_p, c_int, c_int, c_int, c_double, c_void_p, c_int, c_void_p, c_void_p, c_void_p],
    "pyseqm_density_from_evecs": [c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p],
    "pyseqm_mix": [c_void_p, c_int, c_int, c_double, c_void_p, c_void_p, c_void_p, c_void_p],
    "pyseqm_copy_active": [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p],
    "pyseqm_adaptive_mix": [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                            c_void_p, c_void_p, c_void_p],
    "pyseqm_scf_check": [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_double,
                         c_void_p, c_void_p, c_void_p, c_int, c_void_p],
    "pyseqm_gemm": [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p],
    "pyseqm_diis_residual": [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                             c_int, c_int, c_void_p],
    "pyseqm_gemm_bf16x3": [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p],
    "pyseqm_diis_commutator_bf16x3": [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                      c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p],
    "pyseqm_sp2_iterate_many": [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                c_void_p, c_int, c_double, c_int, c_void_p, c_int, c_int, c_void_p],
    "pyseqm_diis_solve": [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p],
    "pyseqm_diis_dots": [c_void_p, c_int, c_void_p, c_longlong, c_int, c_int, c_int, c_void_p, c_void_p, c_int,
                         c_void_p],
    "pyseqm_diis_combine": [c_void_p, c_int, c_int, c_void_p, c_longlong, c_int, c_void_p, c_int, c_void_p,
                            c_void_p],
    "pyseqm_sp2_init": [c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                        c_void_p],
    "pyseqm_sp2_iterate": [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                           c_int, c_double, c_int, c_void_p, c_void_p],
    "pyseqm_dgemm_sym": [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                         c_void_p],
    "pyseqm_sp2_mcweeny": [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                           c_void_p, c_void_p],
    "pyseqm_pair_forces_xl": [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_void_p, c_void_p,
                              c_void_p],
    "pyseqm_xl_energy": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p],
}
# entry points without a layout argument: name -> full argtypes
_RAW_SIGS = {
    "pyseqm_xl_propagate": [c_longlong, c_void_p, c_void_p, c_void_p, c_longlong, c_int, c_void_p, c_double,
                            c_double, c_int, c_void_p],
    "pyseqm_core_repulsion_pairs": [c_longlong, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                    c_void_p],
    "pyseqm_md_verlet": [c_int, c_int, c_double, c_double, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                         c_void_p],
    "pyseqm_build_tiles": [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                           c_void_p, c_void_p, c_void_p, c_void_p],
    "pyseqm_small_solve": [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p],
}
_lib = None


class NativeError(RuntimeError):
    pass


def lib():
    """Load the native library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeError(f"{LIB_PATH} not found: build it with `python -m pyseqm_amd.csrc.build` "
                              "(or __graft_entry__.build()); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = [_P] + args
            f.restype = c_int
        for name, args in _RAW_SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = c_int
        L.pyseqm_fock_scratch_doubles.argtypes = [_P]
        L.pyseqm_fock_scratch_doubles.restype = c_longlong
        L.pyseqm_param_grad_workspace.argtypes = [_P]
        L.pyseqm_param_grad_workspace.restype = c_longlong
        L.pyseqm_eigh_density_workspace.argtypes = [c_int, c_int]
        L.pyseqm_eigh_density_workspace.restype = c_longlong
        L.pyseqm_eigh_density_lds_ok.argtypes = [c_int, c_int]
        L.pyseqm_eigh_density_lds_ok.restype = c_int
        L.pyseqm_fermi_workspace.argtypes = [c_int]
        L.pyseqm_fermi_workspace.restype = c_longlong
        L.pyseqm_ksa_max_rank.argtypes = []
        L.pyseqm_ksa_max_rank.restype = c_int
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS) + list(_RAW_SIGS) + ["pyseqm_fock_scratch_doubles", "pyseqm_param_grad_workspace",
                                            "pyseqm_eigh_density_workspace",
                                            "pyseqm_eigh_density_lds_ok", "pyseqm_fermi_workspace",
                                            "pyseqm_ksa_max_rank"]


def stream_ptr(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


_CODES = {-1: "invalid argument", -2: "kernel launch failed", -3: "unsupported element/method",
          -4: "batch exceeds the library's fixed scratch buffer (PYSEQM_FOCK_SCRATCH_DOUBLES)"}
FOCK_SCRATCH_DOUBLES = 6 * 1024 * 1024  # PYSEQM_FOCK_SCRATCH_DOUBLES of include/pyseqm_hip.h


def call(name, layout, *args):
    rc = getattr(lib(), name)(ctypes.byref(layout), *args)
    if rc != 0:
        raise NativeError(f"{name} failed: {_CODES.get(rc, rc)}")


def call_raw(name, *args):
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        raise NativeError(f"{name} failed: {_CODES.get(rc, rc)}")


MAX_RED_BLOCKS = 256


def red_blocks(nmax):
    """Blocks per molecule of the two-stage reductions (enough to fill the GPU for large n)."""
    return int(min(MAX_RED_BLOCKS, max(1, (nmax * nmax + 8191) // 8192)))


def require_device(device):
    dev = torch.device(device)
    if dev.type != "cuda" or not torch.cuda.is_available():
        raise NativeError("pyseqm_amd runs on a HIP device (MI355X); got device " + str(dev))
    return dev
