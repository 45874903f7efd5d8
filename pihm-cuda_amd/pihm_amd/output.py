"""Host-side mirror of the reference's model output over the HIP library (pihm_output.h).

``read_print_ctrl(para)`` reads the print controls of a ``.para`` file (ReadPrtCtrl,
src/read_func.c:116-160; src/read_para.c:111-180) into ``(kind, interval)`` pairs, and
``HipOutput`` writes the averaged variables as ``<prefix>.<ext>.dat`` / ``.txt`` from sums
kept on the device (UpdPrintVar / PrintData, src/print.c:190-211, 545-603).
"""
import ctypes as C

import numpy as np

from . import _lib

YEARLY, MONTHLY, DAILY, HOURLY = -1, -2, -3, -4
HYDROL_STEP, LS_STEP = 0, 1
# kinds of include/pihm_output.h, in order, with their file extensions
EXTS = (["surf", "unsat", "gw", "stage", "rivgw", "snow", "is", "infil", "recharge", "ec", "ett", "edir"]
        + [f"rivflx{k}" for k in range(11)] + [f"subflx{k}" for k in range(3)]
        + [f"surfflx{k}" for k in range(3)])
KIND = {e: i for i, e in enumerate(EXTS)}
# .para keywords (src/read_para.c:111-180) -> kinds
PARA_KEYS = {"SURF": ["surf"], "UNSAT": ["unsat"], "GW": ["gw"], "RIVSTG": ["stage"], "RIVGW": ["rivgw"],
             "SNOW": ["snow"], "CMC": ["is"], "INFIL": ["infil"], "RECHARGE": ["recharge"], "EC": ["ec"],
             "ETT": ["ett"], "EDIR": ["edir"], "SUBFLX": [f"subflx{k}" for k in range(3)],
             "SURFFLX": [f"surfflx{k}" for k in range(3)]}
PARA_KEYS.update({f"RIVFLX{k}": [f"rivflx{k}"] for k in range(11)})
LS_KINDS = {KIND[e] for e in ("snow", "is", "ec", "ett", "edir")}
RIVER_KINDS = {KIND["stage"], KIND["rivgw"]} | {KIND[f"rivflx{k}"] for k in range(11)}


def parse_interval(s):
    u = s.strip().upper()
    named = {"YEARLY": YEARLY, "MONTHLY": MONTHLY, "DAILY": DAILY, "HOURLY": HOURLY}
    return named[u] if u in named else int(u.split()[0])


def read_print_ctrl(path):
    """[(kind, interval)] for every print control of a .para file that is not 0."""
    out = []
    with open(path) as f:
        for raw in f:
            tok = raw.split("#", 1)[0].split(None, 1)
            if len(tok) == 2 and tok[0].upper() in PARA_KEYS:
                iv = parse_interval(tok[1])
                if iv != 0:
                    out += [(KIND[e], iv) for e in PARA_KEYS[tok[0].upper()]]
    return out


class OutputVar(C.Structure):
    _fields_ = [("kind", C.c_int32), ("intvl", C.c_int32)]


class HipOutput:
    """Output of a single-solve ``HipRHS`` context (and its ``HipLSM``, if any)."""

    def __init__(self, rhs, prefix, variables, ascii=True, lsm=None):
        self.L = _lib.lib()
        self.rhs = rhs
        self.vars = list(variables)
        arr = (OutputVar * max(1, len(self.vars)))(*[OutputVar(k, iv) for k, iv in self.vars])
        h = C.c_void_p()
        rc = self.L.pihm_output_create(rhs.ctx, lsm.h if lsm is not None else None, prefix.encode(), arr,
                                       len(self.vars), 1 if ascii else 0, C.byref(h))
        if rc:
            raise RuntimeError(f"pihm_output_create failed ({rc})")
        self.h = h

    def start(self, y0):
        """The initial state (caller order), which infil's storage change starts from."""
        self._upload(y0)
        rc = self.L.pihm_output_start(self.h, self.rhs.y)
        if rc:
            raise RuntimeError(f"pihm_output_start failed ({rc})")

    def _upload(self, y):
        y = np.ascontiguousarray(y, dtype=np.float64)
        if self.L.pihm_ctx_state_to_device(self.rhs.ctx, y.ctypes.data_as(_lib.c_double_p), self.rhs.y):
            raise RuntimeError("state_to_device failed")

    def update(self, step, y=None):
        """UpdPrintVar; for the model step y (caller order) is uploaded first."""
        if y is not None:
            y = np.ascontiguousarray(y, dtype=np.float64)
            if self.L.pihm_ctx_state_to_device(self.rhs.ctx, y.ctypes.data_as(_lib.c_double_p), self.rhs.y):
                raise RuntimeError("state_to_device failed")
        rc = self.L.pihm_output_update(self.h, int(step), self.rhs.y)
        if rc:
            raise RuntimeError(f"pihm_output_update failed ({rc})")

    def print(self, t, lapse):
        rc = self.L.pihm_output_print(self.h, float(t), float(lapse))
        if rc:
            raise RuntimeError(f"pihm_output_print failed ({rc})")

    def write_ic(self, path, y=None):
        if y is not None:
            y = np.ascontiguousarray(y, dtype=np.float64)
            self.L.pihm_ctx_state_to_device(self.rhs.ctx, y.ctypes.data_as(_lib.c_double_p), self.rhs.y)
        rc = self.L.pihm_output_write_ic(self.h, path.encode(), self.rhs.y)
        if rc:
            raise RuntimeError(f"pihm_output_write_ic failed ({rc})")

    def close(self):
        if getattr(self, "h", None):
            self.L.pihm_output_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_dat(path, n):
    """Records of a .dat file: (times[nrec], values[nrec, n])."""
    raw = np.fromfile(path, dtype=np.float64)
    rec = raw.reshape(-1, n + 1)
    return rec[:, 0], rec[:, 1:]
