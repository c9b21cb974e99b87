"""Device states (posq, correction, fixed-point forces, energy) of several coordinate frames, for GPU tests that drive
the pair and three-body handles through the C ABI as a platform would: one ixh_state per frame."""
import ctypes

import numpy as np

import interax_hip as ih


class Frames:
    def __init__(self, lib, n, positions, box, precision=ih.IXH_MIXED):
        self.lib = lib
        self.n = n
        self.states, self.buffers = [], []
        for pos in positions:
            s = ctypes.c_void_p()
            ih._check(lib.ixh_state_create(0, n, precision, ctypes.byref(s)))
            p = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).reshape(-1))
            ih._check(lib.ixh_state_set_positions(s, ih._dptr(p)))
            if box is not None:
                b = np.ascontiguousarray(np.asarray(box, dtype=np.float64).reshape(-1))
                ih._check(lib.ixh_state_set_box(s, ih._dptr(b)))
            self.states.append(s)
            self.buffers.append(None)
            self.refresh(len(self.states)-1)

    def refresh(self, k):
        buf = ih.Buffers()
        ih._check(self.lib.ixh_state_get_buffers(self.states[k], ctypes.byref(buf)))
        self.buffers[k] = buf

    def share_stream(self):
        """Run every frame's operations on the first frame's stream (one stream, as a platform's Context)."""
        stream = self.buffers[0].stream
        for k, st in enumerate(self.states):
            ih._check(self.lib.ixh_state_set_stream(st, stream))
            self.refresh(k)

    def clear(self, k):
        ih._check(self.lib.ixh_state_clear(self.states[k]))

    def result(self, k, n=None):
        f = np.zeros((n or self.n, 3))
        ih._check(self.lib.ixh_state_get_forces(self.states[k], ih._dptr(f)))
        e = ctypes.c_double()
        ih._check(self.lib.ixh_state_get_energy(self.states[k], ctypes.byref(e)))
        return e.value, f

    def close(self):
        for s in self.states:
            self.lib.ixh_state_destroy(s)
        self.states = []
