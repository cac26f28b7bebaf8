"""Mixed-precision HPD solve with FP64 iterative refinement (no reference
counterpart): the low-precision factor (FP64 diagonal potrf, FP32-stored
panels, FP32, split-BF16x3 or split-BF16x6 trailing update) is refined to the FP64
normwise backward-error target, falling back to GMRES-IR."""
from __future__ import annotations

import ctypes
import dataclasses

from ._lib import IROpts, IRReportC, lib

TRAILING = {"f32": 1, "bf16x3": 2, "bf16x6": 3}
REFINE = {"ir": 0, "cg": 1}


@dataclasses.dataclass
class IROptions:
    nb: int = 128
    trailing: str = "f32"
    max_iter: int = 30
    gmres_restart: int = 20
    force_gmres: bool = False
    allow_shift: bool = False
    tol: float = 1e-14
    refine: str = "ir"  # "ir": classic refinement (GMRES-IR on stagnation), "cg": preconditioned CG
    # finish in plain FP64 when the low-precision route fails (breakdown, a
    # GMRES-IR cycle that does not halve the backward error, max_iter)
    fallback: bool = True

    def c(self) -> IROpts:
        return IROpts(self.nb, TRAILING[self.trailing], self.max_iter, self.gmres_restart if self.gmres_restart else -1,
                      int(self.force_gmres), int(self.allow_shift), self.tol, REFINE[self.refine],
                      int(self.fallback))


@dataclasses.dataclass
class IRReport:
    status: int
    lowprec_info: int
    iterations: int
    gmres_iterations: int
    used_gmres: bool
    shifted: bool
    converged: bool
    backward_error: float
    anorm_inf: float
    fallback: bool = False
    info: int = 0


def workspace_bytes(n: int, k: int, opts: IROptions | None = None) -> int:
    o = (opts or IROptions()).c()
    return int(lib().elemamd_hpd_solve_mixed_workspace(n, k, ctypes.byref(o)))


def HPDSolveMixed(A, B, opts: IROptions | None = None, workspace=None, stream=None) -> IRReport:
    """Solve A X = B (A symmetric positive definite, lower triangle read and
    left unmodified); B is overwritten with X.  ``workspace`` may be a
    preallocated uint8 CUDA tensor of at least ``workspace_bytes`` bytes."""
    import torch

    opts = opts or IROptions()
    n, k = A.Height(), B.Width()
    o = opts.c()
    need = int(lib().elemamd_hpd_solve_mixed_workspace(n, k, ctypes.byref(o)))
    if workspace is None or workspace.numel() < need:
        workspace = torch.empty(need, dtype=torch.uint8, device="cuda")
    rep = IRReportC()
    s = stream if stream is not None else torch.cuda.current_stream()
    rc = lib().elemamd_hpd_solve_mixed_f64(n, k, A.ptr(), A.LDim(), B.ptr(), B.LDim(), ctypes.byref(o),
                                           ctypes.c_void_p(workspace.data_ptr()), workspace.numel(),
                                           ctypes.byref(rep), ctypes.c_void_p(s.cuda_stream))
    if rc not in (0, -3):
        raise RuntimeError(f"elemamd_hpd_solve_mixed_f64 failed with status {rc}")
    return IRReport(rc, rep.lowprec_info, rep.iterations, rep.gmres_iterations, bool(rep.used_gmres),
                    bool(rep.shifted), bool(rep.converged), rep.backward_error, rep.anorm_inf,
                    bool(rep.fallback), rep.info)
