"""ctypes binding of the CPU oracle (oracle/liboracle.so, oracle/liboracle64.so).

Test infrastructure: the oracle is only ever the checker.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_libs = {}


def load(fp64: bool = False) -> C.CDLL:
    name = "liboracle64.so" if fp64 else "liboracle.so"
    if name not in _libs:
        path = os.path.join(ROOT, "oracle", name)
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run 'make -C oracle'")
        L = C.CDLL(path)
        L.oracle_create.restype = C.c_void_p
        L.oracle_create.argtypes = [C.c_void_p]
        L.oracle_destroy.argtypes = [C.c_void_p]
        L.oracle_step.argtypes = [C.c_void_p, C.c_int]
        L.oracle_steps_done.argtypes = [C.c_void_p]
        L.oracle_finalize.argtypes = [C.c_void_p]
        L.oracle_get_field.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int64]
        L.oracle_get_field64.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int64]
        L.oracle_set_field.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int64]
        L.oracle_get_series.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int64]
        L.oracle_get_series64.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int64]
        L.oracle_get_ntff_raw.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int64]
        L.oracle_get_ff_time.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
        L.oracle_get_video.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int64]
        L.oracle_port_exchange.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        _libs[name] = L
    return _libs[name]


SERIES = {"port_v": 0, "port_i": 1, "incident": 2, "e_theta": 3, "e_phi": 4, "reflected": 5,
          "transmitted": 6, "ff_pc": 7, "spice_v": 20, "spice_i": 21, "reflected_tm": 22, "transmitted_tm": 23}
IMAG = 0x100    # the imaginary part of a kmax plan's series or field
FIELDS = {"Ex": 0, "Ey": 1, "Ez": 2, "Hx": 3, "Hy": 4, "Hz": 5}


class Oracle:
    def __init__(self, plan, fp64: bool = False):
        self.L = load(fp64)
        self.plan = plan
        self.h = self.L.oracle_create(plan.view_ptr)
        if not self.h:
            raise RuntimeError("oracle_create failed")

    def step(self, n: int) -> None:
        if self.L.oracle_step(self.h, int(n)) != 0:
            raise RuntimeError("oracle: step count exceeds time_steps")

    def finalize(self) -> None:
        self.L.oracle_finalize(self.h)

    def port_exchange(self, v_in, n_spice: int):
        """One SPICE exchange (see aem_port_exchange): finish the waiting step
        with node voltages v_in (None for the first call), then, unless
        n_spice is 0, run the next step up to its gap currents and return
        them."""
        v = None if v_in is None else np.ascontiguousarray(v_in, np.float32)
        i = np.zeros(max(n_spice, 1), np.float32)
        if self.L.oracle_port_exchange(self.h, None if v is None else v.ctypes.data,
                                       i.ctypes.data if n_spice else None) != 0:
            raise RuntimeError("oracle: SPICE exchange out of order")
        return i[:n_spice] if n_spice else None

    def series(self, kind: str, idx: int = 0, n: int = None, imag: bool = False) -> np.ndarray:
        n = self.plan.time_steps if n is None else n
        out = np.zeros(n, np.float32)
        if self.L.oracle_get_series(self.h, SERIES[kind] | (IMAG if imag else 0), idx, out.ctypes.data, n) != 0:
            raise ValueError(kind)
        return out

    def series64(self, kind: str, idx: int = 0) -> np.ndarray:
        """A port or incident series in double (the FP64 build's precision)."""
        out = np.zeros(self.plan.time_steps, np.float64)
        if self.L.oracle_get_series64(self.h, SERIES[kind], idx, out.ctypes.data, out.size) != 0:
            raise ValueError(kind)
        return out

    def cseries(self, kind: str, idx: int = 0) -> np.ndarray:
        """A kmax plan's complex series as complex64."""
        return self.series(kind, idx) + 1j * self.series(kind, idx, imag=True).astype(np.complex64)

    def field(self, name: str, imag: bool = False) -> np.ndarray:
        v = self.plan.view
        out = np.zeros(v.ncell, np.float32)
        if self.L.oracle_get_field(self.h, FIELDS[name] | (IMAG if imag else 0), out.ctypes.data, v.ncell) != 0:
            raise ValueError(name)
        return out.reshape(self.plan.shape)

    def field64(self, name: str) -> np.ndarray:
        """The component in double: full precision of the FP64 build."""
        v = self.plan.view
        out = np.zeros(v.ncell, np.float64)
        self.L.oracle_get_field64(self.h, FIELDS[name], out.ctypes.data, v.ncell)
        return out.reshape(self.plan.shape)

    def set_field(self, name: str, values, imag: bool = False) -> None:
        a = np.ascontiguousarray(values, np.float32).reshape(-1)
        if self.L.oracle_set_field(self.h, FIELDS[name] | (IMAG if imag else 0), a.ctypes.data, a.size) != 0:
            raise ValueError(name)

    def ntff_raw(self, which: int) -> np.ndarray:
        v = self.plan.view
        out = np.zeros(v.n_angles * v.len_ff, np.float64)
        self.L.oracle_get_ntff_raw(self.h, which, out.ctypes.data, out.size)
        return out.reshape(v.n_angles, v.len_ff)

    def ff_time(self) -> np.ndarray:
        n = self.plan.view.n_angles
        out = np.zeros(n, np.float32)
        self.L.oracle_get_ff_time(self.h, out.ctypes.data, n)
        return out

    def video(self, comp: int, step: int) -> np.ndarray:
        v = self.plan.view
        out = np.zeros(v.vid_size1 * v.vid_size2, np.float32)
        if self.L.oracle_get_video(self.h, comp, step, out.ctypes.data, out.size) != 0:
            raise ValueError("video")
        return out.reshape(v.vid_size2, v.vid_size1)

    def __del__(self):
        try:
            if self.h:
                self.L.oracle_destroy(self.h)
        except Exception:
            pass
