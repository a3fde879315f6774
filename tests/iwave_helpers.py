"""Helpers for interaction-field runs shared by the oracle and GPU tests."""
import ctypes as C

import numpy as np

from oceanlib_amd import capi


def idesc(lib_init, **kw) -> capi.InteractionDesc:
    d = capi.InteractionDesc()
    lib_init(C.byref(d))
    for k, v in kw.items():
        setattr(d, k, v)
    return d


def disturbance(rec) -> capi.Disturbance:
    d = capi.Disturbance()
    for k, v in rec.items():
        setattr(d, k, v)
    return d


class Field:
    """One Python interface over the three implementations' C functions.
    `kind` is 'orc' (oracle), 'oref' (reference build) or 'gpu' (library)."""

    def __init__(self, kind, lib, desc):
        self.kind, self.lib = kind, lib
        self.n = desc.size
        if kind == "orc":
            self.h = lib.orc_iw_create(C.byref(desc))
        else:
            st = C.c_int(-1)
            fn = lib.oref_interaction_create if kind == "oref" else lib.ocean_interaction_create
            self.h = fn(C.byref(desc), C.byref(st))
        assert self.h, f"{kind} create failed"

    def _f(self, name):
        pre = {"orc": "orc_iw_", "oref": "oref_interaction_", "gpu": "ocean_interaction_"}
        return getattr(self.lib, pre[self.kind] + name)

    def add(self, rec):
        d = disturbance(rec)
        self._f("add")(self.h, C.byref(d))

    def update(self, dt):
        self._f("update")(self.h, np.float32(dt))

    def recenter(self, x, z):
        return self._f("recenter")(self.h, np.float32(x), np.float32(z))

    def set_obstruction(self, mask):
        if mask is None:
            self._f("set_obstruction")(self.h, None)
        else:
            m = np.ascontiguousarray(mask, dtype=np.uint8)
            self._f("set_obstruction")(self.h, m.ctypes.data_as(C.POINTER(C.c_uint8)))

    def field(self):
        if self.kind == "orc":
            p = self.lib.orc_iw_field(self.h)
        else:
            p = self._f("get_buffers")(self.h).field
        return np.ctypeslib.as_array(p, shape=(4 * self.n * self.n,)).copy()

    def sample(self, x, z):
        s = capi.InteractionSample()
        self._f("sample_at")(self.h, np.float32(x), np.float32(z), C.byref(s))
        return s

    def dropped(self):
        return self._f("dropped")(self.h)

    def energy(self) -> bytes:
        """InteractionField::energy() as its 8 bytes (bit-for-bit comparison)."""
        fn = self.lib.ocean_amd_interaction_energy if self.kind == "gpu" else self._f("energy")
        return np.float64(fn(self.h)).tobytes()

    def fixed_dt(self):
        return self._f("fixed_dt")(self.h)

    def dt_limit(self):
        return self._f("dt_limit")(self.h)

    def close(self):
        if self.h:
            self._f("destroy")(self.h)
            self.h = None
