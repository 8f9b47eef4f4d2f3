"""The step before the factorization, on one MI355X: the solid element matrices of a mesh added
in element order into the upper triangle by columns of K_ff, M_ff or K_ff - shift M_ff, in device
memory, where ``Factor.refactor_device`` reads the values.

Mirrors the reference's assembly chain (FieldCraft 0.8.1, crates/solver): ``assemble``
(analysis.rs:199-233) with ``add_block`` / ``add_scalar`` (sparse.rs:199-243), ``Reduction::reduce``
for a plain selection (constraints.rs:820-906), ``Csr::add_scaled`` (sparse.rs:81-87) and
``UpperColumns::from_symmetric`` (cholesky.rs:110-144), over the element kernels of crates/fem
(8-node hexahedra with full integration, 10-node tetrahedra with their consistent mass, isotropic
materials). The values are the reference's bit for bit. A degenerate element raises
``DegenerateElement`` naming the first one in element order, as the reference's sequential loop
does. Nothing here computes on the CPU; a missing library raises.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import cholesky as _ch

HEX8, TET10 = 0, 1
ELEMENT_KINDS = {"hex8": HEX8, "tet10": TET10}
K, M, SHIFTED = 0, 1, 2
WHAT = {"K": K, "M": M, "shifted": SHIFTED}

_configured = False


def _lib():
    global _configured
    L = _ch.lib()
    if not _configured:
        vp, ep = C.c_void_p, C.POINTER(_ch._Error)
        f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
        u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
        L.fc_assembly_new.argtypes = [C.c_int32, C.c_uint64, f64p, C.c_uint64, i64p, f64p, C.c_uint64, C.c_double,
                                      C.c_double, C.c_uint64, i64p, u64p, u64p, C.c_int32, C.c_int32,
                                      C.POINTER(vp), ep]
        L.fc_assemble_device.argtypes = [vp, C.c_int32, C.c_double, vp, C.c_uint64, vp, ep]
        L.fc_assembly_info.argtypes = [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_double),
                                       C.POINTER(C.c_double)]
        L.fc_assembly_free.argtypes = [vp]
        L.fc_assembly_free.restype = None
        _configured = True
    return L


class Assembly:
    """A mesh of one solid element kind, its free equations and the upper-column pattern of K_ff
    (what ``Symbolic`` was built from), planned once; ``assemble`` then writes K_ff, M_ff or the
    shifted matrix's values into device memory."""

    def __init__(self, kind: str, xyz, conn, young, poisson: float, eq_dof, ptr, idx, *, density: float = 0.0,
                 device: int = 0, timing: bool = False):
        L = _lib()
        xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
        conn = np.ascontiguousarray(conn, np.int64)
        young = np.ascontiguousarray(np.atleast_1d(young), np.float64)
        eq_dof = np.ascontiguousarray(eq_dof, np.int64)
        self.nnz = int(np.asarray(ptr)[-1])
        self.device = device
        self._L = L  # kept for __del__: module globals may be gone at interpreter exit
        self._h = C.c_void_p()
        err = _ch._Error()
        st = L.fc_assembly_new(ELEMENT_KINDS[kind], len(xyz), xyz.reshape(-1) if xyz.size else np.zeros(1),
                               len(conn), conn.reshape(-1) if conn.size else np.zeros(1, np.int64), young, len(young),
                               float(poisson), float(density), len(eq_dof),
                               eq_dof if eq_dof.size else np.zeros(1, np.int64), _ch._u64(ptr), _ch._u64(idx),
                               int(device), int(bool(timing)), C.byref(self._h), C.byref(err))
        _ch._raise(st, err)

    @staticmethod
    def of_problem(p, upper_ptr, upper_idx, *, device: int = 0, timing: bool = False) -> "Assembly":
        """From a generated problem (``problems.block``): its mesh, materials and equations."""
        return Assembly(p.kind, p.xyz, p.conn, p.young, p.poisson, p.eq_dof, upper_ptr, upper_idx,
                        density=p.rho, device=device, timing=timing)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._L.fc_assembly_free(h)
            self._h = None

    def assemble_device(self, out_ptr: int, what: str = "K", shift: float = 0.0, stream: int | None = None) -> None:
        """Write the values of ``what`` ("K", "M" or "shifted" = K + (-shift) M) into ``out_ptr``
        (device memory of ``nnz`` doubles); waits for the result."""
        err = _ch._Error()
        st = _lib().fc_assemble_device(self._h, WHAT[what], float(shift), C.c_void_p(out_ptr), self.nnz,
                                       C.c_void_p(stream) if stream else None, C.byref(err))
        _ch._raise(st, err)

    def assemble(self, what: str = "K", shift: float = 0.0):
        """``assemble_device`` into a new torch tensor on the assembly's device."""
        import torch

        out = torch.empty(max(self.nnz, 1), dtype=torch.float64, device=f"cuda:{self.device}")
        self.assemble_device(out.data_ptr(), what, shift, torch.cuda.current_stream(self.device).cuda_stream)
        return out[:self.nnz]

    def info(self) -> dict:
        ck, cm = C.c_uint64(), C.c_uint64()
        te, tg = C.c_double(), C.c_double()
        _lib().fc_assembly_info(self._h, C.byref(ck), C.byref(cm), C.byref(te), C.byref(tg))
        return dict(contributions_k=int(ck.value), contributions_m=int(cm.value), elements_s=te.value,
                    gather_s=tg.value)
