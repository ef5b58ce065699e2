"""ctypes wrapper of the CPU oracle (oracle/build/libdtb_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the timed CPU baseline.
The product never imports it.
"""
import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_float, c_int32, c_int64, c_uint8, c_uint32, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libdtb_oracle.so")
NCOL = 22


class _Net(ctypes.Structure):
    _fields_ = [
        ("n", c_int64), ("npop", c_int32), ("pop_begin", c_void_p), ("props", c_void_p), ("rowptr", c_void_p),
        ("col", c_void_p), ("kind", c_void_p), ("w", c_void_p), ("dt", c_float), ("t_steps", c_int32),
        ("seed", c_uint64), ("ou_on", c_void_p), ("ou_mean", c_void_p), ("ou_sigma", c_void_p),
        ("ou_tau", c_void_p),
    ]


class _State32(ctypes.Structure):
    _fields_ = [("V", c_void_p), ("t", c_void_p), ("f", c_void_p), ("i_ou", c_void_p), ("J", c_void_p),
                ("I_syn", c_void_p), ("step", c_int64)]


_State64 = _State32

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    lib.dtbo_philox4x32_10.argtypes = [POINTER(c_uint32), POINTER(c_uint32), POINTER(c_uint32)]
    lib.dtbo_philox4x32_10.restype = None
    lib.dtbo_log.argtypes = [c_float]
    lib.dtbo_log.restype = c_float
    lib.dtbo_sin_halfpi.argtypes = [c_float]
    lib.dtbo_sin_halfpi.restype = c_float
    lib.dtbo_cos_halfpi.argtypes = [c_float]
    lib.dtbo_cos_halfpi.restype = c_float
    lib.dtbo_normal_from_bits.argtypes = [c_uint32, c_uint32]
    lib.dtbo_normal_from_bits.restype = c_float
    lib.dtbo_uniform_from_bits.argtypes = [c_uint32]
    lib.dtbo_uniform_from_bits.restype = c_float
    lib.dtbo_ou_normal.argtypes = [c_uint64, c_uint32, c_uint64]
    lib.dtbo_ou_normal.restype = c_float
    lib.dtbo_normals.argtypes = [c_uint64, c_uint32, c_uint64, c_int64, c_void_p]
    lib.dtbo_normals.restype = None
    lib.dtbo_init32.argtypes = [POINTER(_Net), POINTER(_State32)]
    lib.dtbo_init64.argtypes = [POINTER(_Net), POINTER(_State64)]
    lib.dtbo_run32.argtypes = [POINTER(_Net), POINTER(_State32), c_int32, c_int32, c_void_p, c_void_p, c_void_p,
                               c_void_p, c_int32]
    lib.dtbo_stat_q.argtypes = [c_float]
    lib.dtbo_stat_q.restype = c_int64
    lib.dtbo_run32.restype = c_int32
    lib.dtbo_run64.argtypes = [POINTER(_Net), POINTER(_State64), c_int32, c_int32, c_void_p, c_void_p, c_void_p,
                               c_void_p, c_int32]
    lib.dtbo_run64.restype = c_int32
    lib.dtbo_dense_delta.argtypes = [POINTER(_Net), c_void_p, c_void_p]
    lib.dtbo_scatter_q40.argtypes = [POINTER(_Net), c_void_p, c_void_p]
    _lib = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(c_void_p)


def philox(ctr, key):
    lib = load()
    c = (c_uint32 * 4)(*[int(x) & 0xffffffff for x in ctr])
    k = (c_uint32 * 2)(*[int(x) & 0xffffffff for x in key])
    o = (c_uint32 * 4)()
    lib.dtbo_philox4x32_10(c, k, o)
    return [int(x) for x in o]


def normals(seed, neuron0, substep, count):
    out = np.empty(count, dtype=np.float32)
    load().dtbo_normals(seed, neuron0, substep, count, _p(out))
    return out


class OracleNet:
    """Holds numpy arrays of a network plus OU settings in oracle form."""

    def __init__(self, props, rowptr, col, w, kind_u8, dt=0.1, t_steps=1, seed=0):
        self.props = np.ascontiguousarray(props, dtype=np.float32)
        self.n = self.props.shape[0]
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
        self.col = np.ascontiguousarray(col, dtype=np.uint32)
        self.w = np.ascontiguousarray(w, dtype=np.float32).reshape(-1, 2)
        self.kind = np.ascontiguousarray(kind_u8, dtype=np.uint8)
        sub = self.props[:, 3].astype(np.int64)
        starts = np.flatnonzero(np.r_[True, sub[1:] != sub[:-1]]) if self.n else np.zeros(0, np.int64)
        self.pop_ids = sub[starts].astype(np.uint32)
        self.pop_begin = np.r_[starts, self.n].astype(np.int64)
        self.npop = len(self.pop_ids)
        self.dt = float(dt)
        self.t_steps = int(t_steps)
        self.seed = int(seed)
        self.ou_on = np.zeros(self.npop, dtype=np.uint8)
        self.ou_mean = np.zeros(self.npop, dtype=np.float32)
        self.ou_sigma = np.zeros(self.npop, dtype=np.float32)
        self.ou_tau = np.ones(self.npop, dtype=np.float32)

    @staticmethod
    def from_network(net, dt=0.1, t_steps=1, seed=0):
        """net: NetworkArrays; uint8 weight codes u are decoded as
        float32(u) / float32(255) (presynaptic_voltage.cpp:171-176), float16
        weights converted to float32 (:165-169)."""
        w = net.w
        if w.dtype == np.uint8:
            w = w.astype(np.float32) / np.float32(255.0)
        return OracleNet(net.props, net.rowptr, net.col, w.astype(np.float32), net.kind_per_synapse(), dt, t_steps,
                         seed)

    def set_ou(self, brain_ids, mean, sigma, tau):
        ids = np.atleast_1d(np.asarray(brain_ids))
        mean = np.broadcast_to(np.asarray(mean, np.float32), ids.shape)
        sigma = np.broadcast_to(np.asarray(sigma, np.float32), ids.shape)
        tau = np.broadcast_to(np.asarray(tau, np.float32), ids.shape)
        for k, b in enumerate(ids):
            idx = np.flatnonzero(self.pop_ids == b)
            if idx.size:
                q = idx[0]
                self.ou_on[q] = 1
                self.ou_mean[q] = mean[k]
                self.ou_sigma[q] = sigma[k]
                self.ou_tau[q] = tau[k]

    def struct(self):
        return _Net(self.n, self.npop, _p(self.pop_begin), _p(self.props), _p(self.rowptr), _p(self.col),
                    _p(self.kind), _p(self.w), self.dt, self.t_steps, self.seed, _p(self.ou_on), _p(self.ou_mean),
                    _p(self.ou_sigma), _p(self.ou_tau))


class State:
    def __init__(self, n, dtype):
        self.V = np.zeros(n, dtype=dtype)
        self.t = np.zeros(n, dtype=np.int32)
        self.f = np.zeros(n, dtype=np.uint8)
        self.i_ou = np.zeros(n, dtype=dtype)
        self.J = np.zeros((n, 4), dtype=dtype)
        self.I_syn = np.zeros(n, dtype=dtype)
        self.step = 0

    def struct(self):
        return _State32(_p(self.V), _p(self.t), _p(self.f), _p(self.i_ou), _p(self.J), _p(self.I_syn),
                        self.step)


class Sim:
    """Oracle simulation of one network in fp32 (canonical order) or fp64."""

    def __init__(self, net, precision=32, nthreads=0):
        self.net = net
        self.prec = precision
        self.nthreads = nthreads
        self.st = State(net.n, np.float32 if precision == 32 else np.float64)
        self._ns = net.struct()
        s = self.st.struct()
        if precision == 32:
            load().dtbo_init32(ctypes.byref(self._ns), ctypes.byref(s))
        else:
            load().dtbo_init64(ctypes.byref(self._ns), ctypes.byref(s))
        self.st.step = s.step

    def load_state(self, V, t, f, i_ou, J, I_syn, step):
        """Starts from a given reference-visible state (e.g. one read back
        from the GPU with dtb_read_state)."""
        dt = self.st.V.dtype
        self.st.V[:] = np.asarray(V, dtype=dt)
        self.st.t[:] = np.asarray(t, dtype=np.int32)
        self.st.f[:] = np.asarray(f, dtype=np.uint8)
        self.st.i_ou[:] = np.asarray(i_ou, dtype=dt)
        self.st.J[:] = np.asarray(J, dtype=dt).reshape(-1, 4)
        self.st.I_syn[:] = np.asarray(I_syn, dtype=dt)
        self.st.step = int(step)

    def refresh(self):
        """Re-reads network parameters (after OU or property changes)."""
        self._ns = self.net.struct()

    def reset_ou_current(self):
        """Sets i_ou to the OU mean of every configured population."""
        for q in range(self.net.npop):
            if self.net.ou_on[q]:
                self.st.i_ou[self.net.pop_begin[q]:self.net.pop_begin[q + 1]] = self.net.ou_mean[q]

    def run(self, nsteps, use_ou=True, raster=False, vtrace=False, pop_counts=False, force_raster=None,
            margins=False, stats=False):
        """stats: also returns out["stat_sums"] [nsteps][npop][6] (exact sums of
        V, I_syn, AMPA, NMDA, GABA_A, GABA_B currents as multiples of 2^-24)
        and out["stat_means"] (the same divided by the non-blocked counts)."""
        n = self.net.n
        out = {}
        r = np.zeros((nsteps, n), dtype=np.uint8) if raster else None
        vdt = np.float32 if self.prec == 32 else np.float64
        v = np.zeros((nsteps, n), dtype=vdt) if vtrace else None
        s = self.st.struct()
        if self.prec == 32:
            pc = np.zeros((nsteps, self.net.npop), dtype=np.uint32) if pop_counts else None
            ss = np.zeros((nsteps, self.net.npop, 6), dtype=np.int64) if stats else None
            rc = load().dtbo_run32(ctypes.byref(self._ns), ctypes.byref(s), nsteps, int(use_ou), _p(r), _p(v),
                                   _p(pc), _p(ss), self.nthreads)
            if pc is not None:
                out["freq"] = pc
            if ss is not None:
                out["stat_sums"] = ss
                out["stat_means"] = stat_means(ss, self.net.props, self.net.pop_begin)
        else:
            fr = None if force_raster is None else np.ascontiguousarray(force_raster, dtype=np.uint8)
            m = np.zeros(nsteps, dtype=np.float64) if margins else None
            rc = load().dtbo_run64(ctypes.byref(self._ns), ctypes.byref(s), nsteps, int(use_ou), _p(fr), _p(r),
                                   _p(v), _p(m), self.nthreads)
            if m is not None:
                out["margin"] = m
        if rc != 0:
            raise ValueError("oracle rejected the run configuration")
        self.st.step = s.step
        if r is not None:
            out["raster"] = r
        if v is not None:
            out["V"] = v
        return out


def stat_means(sums, props, pop_begin):
    """Per-population means from exact 2^-24 sums: float32 of
    (double) sum * 2^-24 / (neurons of the population that are not blocked)."""
    blocked = props[:, 1] != 0
    cnt = np.array([np.count_nonzero(~blocked[pop_begin[q]:pop_begin[q + 1]]) for q in range(len(pop_begin) - 1)],
                   dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        return (sums.astype(np.float64) * 2.0 ** -24 / cnt[None, :, None]).astype(np.float32)


def dense_delta(net, fired):
    d = np.zeros((net.n, 4), dtype=np.float64)
    ns = net.struct()
    f = np.ascontiguousarray(fired, dtype=np.uint8)
    load().dtbo_dense_delta(ctypes.byref(ns), _p(f), _p(d))
    return d


def scatter_q40(net, fired):
    d = np.zeros((net.n, 4), dtype=np.int64)
    ns = net.struct()
    f = np.ascontiguousarray(fired, dtype=np.uint8)
    load().dtbo_scatter_q40(ctypes.byref(ns), _p(f), _p(d))
    return d
