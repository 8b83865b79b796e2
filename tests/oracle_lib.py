"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the tests.

Test infrastructure only: the oracle is the checker, never the thing under
test.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_PATH = ROOT / "oracle" / "liboracle.so"

_F = ctypes.POINTER(ctypes.c_float)
_D = ctypes.POINTER(ctypes.c_double)
_U = ctypes.POINTER(ctypes.c_uint32)


class OracleProblem(ctypes.Structure):
    _fields_ = [
        ("D", ctypes.c_uint32), ("C", ctypes.c_uint32), ("H", ctypes.c_uint32),
        ("K", ctypes.c_uint32), ("lambda_", ctypes.c_float), ("cov_full", ctypes.c_uint32),
        ("sigma", _F), ("u_min", _F), ("u_max", _F), ("seed", ctypes.c_uint64),
        ("model", ctypes.c_char_p), ("dynamics", ctypes.c_char_p),
        ("stage_cost", ctypes.c_char_p), ("terminal_cost", ctypes.c_char_p),
        ("model_params", ctypes.c_void_p), ("cost_params", ctypes.c_void_p),
        ("noise_f16", ctypes.c_uint32),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not ORACLE_PATH.exists():
            raise ImportError(f"{ORACLE_PATH} missing: run `make -C oracle`")
        L = ctypes.CDLL(str(ORACLE_PATH))
        P = ctypes.POINTER(OracleProblem)
        sig = {
            "oracle_philox4x32": (None, [_U, _U, _U]),
            "oracle_library_philox4x32": (None, [_U, _U, _U]),
            "oracle_library_philox4x32_split": (None, [_U, _U, _U]),
            "oracle_philox2x32": (None, [_U, ctypes.c_uint32, _U]),
            "oracle_normals4": (None, [ctypes.c_uint32] * 5 + [ctypes.c_uint64, _F]),
            "oracle_normals4_64": (None, [ctypes.c_uint32] * 5 + [ctypes.c_uint64, _D]),
            "oracle_mathf": (None, [ctypes.c_int, _F, _F, ctypes.c_size_t]),
            "oracle_fmodf": (None, [_F, _F, _F, ctypes.c_size_t]),
            "oracle_math2f": (None, [ctypes.c_int, _F, _F, _F, ctypes.c_size_t]),
            "oracle_rollout": (ctypes.c_int, [P, ctypes.c_uint32, ctypes.c_uint32, _F, _F, _F, ctypes.c_int]),
            "oracle_rollout_range": (ctypes.c_int, [P, ctypes.c_uint32, ctypes.c_uint32, _F, _F,
                                                    ctypes.c_uint32, ctypes.c_uint32, _F]),
            "oracle_rollout64": (ctypes.c_int, [P, ctypes.c_uint32, ctypes.c_uint32, _F, _D, _D, ctypes.c_int]),
            "oracle_weights": (ctypes.c_int, [P, _F, _F, _F]),
            "oracle_weights64": (ctypes.c_int, [P, _D, _D, _D]),
            "oracle_reduce": (ctypes.c_int, [P, ctypes.c_uint32, ctypes.c_uint32, _F, _F, _F]),
            "oracle_reduce64": (ctypes.c_int, [P, ctypes.c_uint32, ctypes.c_uint32, _D, _D, _D]),
            "oracle_reduce_abs": (ctypes.c_int, [P, ctypes.c_uint32, ctypes.c_uint32, _D, _D]),
            "oracle_step": (ctypes.c_int, [P, ctypes.c_uint32, ctypes.c_uint32, _F, _F, _F, _F, _F, _F,
                                           ctypes.c_int]),
            "oracle_step64": (ctypes.c_int, [P, ctypes.c_uint32, ctypes.c_uint32, _F, _D, _D, _D, _D, _D,
                                             ctypes.c_int]),
            "oracle_simulate": (ctypes.c_int, [P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _F, _F,
                                               _F, _F, _F, _F, ctypes.c_int]),
            "oracle_simulate64": (ctypes.c_int, [P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _F,
                                                 _F, _F, _F, _F, _F, ctypes.c_int]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a, kind=_F):
    if a is None:
        return None
    return a.ctypes.data_as(kind)


class Problem:
    """Keeps the arrays and blobs an OracleProblem points to alive."""

    def __init__(self, model, D, C, H, K, lam, sigma, u_min, u_max, seed, dynamics="mppi_dynamics",
                 stage_cost="", terminal_cost="", model_params=None, cost_params=None,
                 cov_full=False, noise_f16=False):
        self.D, self.C, self.H, self.K = D, C, H, K
        self.sigma = np.ascontiguousarray(sigma, np.float32)
        self.u_min = np.ascontiguousarray(u_min, np.float32)
        self.u_max = np.ascontiguousarray(u_max, np.float32)
        self._mp = ctypes.create_string_buffer(bytes(model_params), len(model_params)) if model_params else None
        self._cp = ctypes.create_string_buffer(bytes(cost_params), len(cost_params)) if cost_params else None
        self._names = [s.encode() for s in (model, dynamics, stage_cost or "", terminal_cost or "")]
        self.s = OracleProblem(
            D, C, H, K, lam, 1 if cov_full else 0, _p(self.sigma), _p(self.u_min), _p(self.u_max),
            seed, self._names[0], self._names[1], self._names[2], self._names[3],
            ctypes.cast(self._mp, ctypes.c_void_p) if self._mp else None,
            ctypes.cast(self._cp, ctypes.c_void_p) if self._cp else None, 1 if noise_f16 else 0)

    @property
    def ref(self):
        return ctypes.byref(self.s)

    def rollout(self, step, x0, u_nom, agent=0, threads=1, fp64=False):
        x0 = np.ascontiguousarray(x0, np.float32)
        if fp64:
            un = np.ascontiguousarray(u_nom, np.float64).ravel()
            out = np.empty(self.K, np.float64)
            rc = lib().oracle_rollout64(self.ref, step, agent, _p(x0), _p(un, _D), _p(out, _D), threads)
        else:
            un = np.ascontiguousarray(u_nom, np.float32).ravel()
            out = np.empty(self.K, np.float32)
            rc = lib().oracle_rollout(self.ref, step, agent, _p(x0), _p(un), _p(out), threads)
        assert rc == 0
        return out

    def rollout_range(self, step, x0, u_nom, k0, k1, agent=0):
        x0 = np.ascontiguousarray(x0, np.float32)
        un = np.ascontiguousarray(u_nom, np.float32).ravel()
        out = np.full(self.K, np.nan, np.float32)
        assert lib().oracle_rollout_range(self.ref, step, agent, _p(x0), _p(un), k0, k1, _p(out)) == 0
        return out[k0:k1]

    def rollout_surrogate_range(self, step, x0, u_nom, k0, k1, params, norm80, agent=0):
        """Costs of samples [k0, k1) with the surrogate network as dynamics."""
        x0 = np.ascontiguousarray(x0, np.float32)
        un = np.ascontiguousarray(u_nom, np.float32).ravel()
        P = np.ascontiguousarray(params, np.float32)
        nb = np.ascontiguousarray(norm80, np.float32)
        out = np.full(self.K, np.nan, np.float32)
        L = lib()
        L.oracle_rollout_surrogate_range.restype = ctypes.c_int
        assert L.oracle_rollout_surrogate_range(self.ref, ctypes.c_uint32(step), ctypes.c_uint32(agent), _p(x0),
                                                _p(un), ctypes.c_uint32(k0), ctypes.c_uint32(k1), _p(P),
                                                _p(nb), _p(out)) == 0
        return out[k0:k1]

    def stage_costs(self, states, controls, steps):
        """stage(x_i, u_i, time {0, steps_i}) for each row."""
        s = np.ascontiguousarray(states, np.float32).reshape(-1, self.D)
        u = np.ascontiguousarray(controls, np.float32).reshape(-1, self.C)
        st = np.ascontiguousarray(steps, np.uint32)
        out = np.empty(s.shape[0], np.float32)
        L = lib()
        L.oracle_stage_costs.restype = ctypes.c_int
        assert L.oracle_stage_costs(self.ref, ctypes.c_uint32(s.shape[0]), _p(s), _p(u),
                                    st.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), _p(out)) == 0
        return out

    def closed_loop_cost(self, states, controls, terminal, start_step=0):
        """Realised cost of a closed loop from simulate(): each step's stage
        cost on the state the plant reached with the applied control."""
        st = np.asarray(states, np.float32).reshape(-1, self.D)
        nxt = np.concatenate([st[1:], np.asarray(terminal, np.float32).reshape(1, self.D)])
        steps = np.arange(start_step, start_step + st.shape[0], dtype=np.uint32)
        return self.stage_costs(nxt, controls, steps)

    def weights(self, costs, fp64=False):
        if fp64:
            c = np.ascontiguousarray(costs, np.float64)
            w, d = np.empty(self.K, np.float64), np.empty(3, np.float64)
            lib().oracle_weights64(self.ref, _p(c, _D), _p(w, _D), _p(d, _D))
        else:
            c = np.ascontiguousarray(costs, np.float32)
            w, d = np.empty(self.K, np.float32), np.empty(3, np.float32)
            lib().oracle_weights(self.ref, _p(c), _p(w), _p(d))
        return w, d

    def reduce(self, step, w, u_nom, agent=0, fp64=False):
        if fp64:
            w = np.ascontiguousarray(w, np.float64)
            un = np.ascontiguousarray(u_nom, np.float64).ravel()
            out = np.empty(self.H * self.C, np.float64)
            assert lib().oracle_reduce64(self.ref, step, agent, _p(w, _D), _p(un, _D), _p(out, _D)) == 0
        else:
            w = np.ascontiguousarray(w, np.float32)
            un = np.ascontiguousarray(u_nom, np.float32).ravel()
            out = np.empty(self.H * self.C, np.float32)
            assert lib().oracle_reduce(self.ref, step, agent, _p(w), _p(un), _p(out)) == 0
        return out.reshape(self.H, self.C)

    def reduce_abs(self, step, w, agent=0):
        """sum_k w_k |eps_k| per control [H, C] in FP64: the magnitude of the
        update's sum."""
        w = np.ascontiguousarray(w, np.float64)
        out = np.empty(self.H * self.C, np.float64)
        assert lib().oracle_reduce_abs(self.ref, step, agent, _p(w, _D), _p(out, _D)) == 0
        return out.reshape(self.H, self.C)

    def step(self, step, x0, u_nom, agent=0, threads=1, fp64=False):
        """Returns (u_next [H, C], diag [3], costs [K], weights [K])."""
        x0 = np.ascontiguousarray(x0, np.float32)
        dt, kind = (np.float64, _D) if fp64 else (np.float32, _F)
        un = np.ascontiguousarray(u_nom, dt).ravel()
        u = np.empty(self.H * self.C, dt)
        d = np.empty(3, dt)
        c = np.empty(self.K, dt)
        w = np.empty(self.K, dt)
        fn = lib().oracle_step64 if fp64 else lib().oracle_step
        assert fn(self.ref, step, agent, _p(x0), _p(un, kind), _p(u, kind), _p(d, kind), _p(c, kind),
                  _p(w, kind), threads) == 0
        return u.reshape(self.H, self.C), d, c, w

    def simulate(self, T, start_step, x0_packed, n_agents=1, u_nom_packed=None, threads=1, fp64=False):
        x0 = np.ascontiguousarray(x0_packed, np.float32).ravel()
        un = None if u_nom_packed is None else np.ascontiguousarray(u_nom_packed, np.float32).ravel()
        st = np.empty((T, n_agents, self.D), np.float32)
        ct = np.empty((T, n_agents, self.C), np.float32)
        kt = np.empty((T, n_agents), np.float32)
        tt = np.empty((n_agents, self.D), np.float32)
        fn = lib().oracle_simulate64 if fp64 else lib().oracle_simulate
        assert fn(self.ref, T, start_step, n_agents, _p(x0), _p(un), _p(st), _p(ct), _p(kt), _p(tt),
                  threads) == 0
        return st, ct, kt, tt


def _prefilter(fn, cost_blob, xy):
    blob = np.frombuffer(bytes(cost_blob), np.uint8).copy()
    pts = np.ascontiguousarray(xy, np.float32).reshape(-1, 2)
    out = np.empty(pts.shape[0], np.uint32)
    f = getattr(lib(), fn)
    f.restype = ctypes.c_int
    assert f(blob.ctypes.data_as(ctypes.c_void_p), ctypes.c_uint32(pts.shape[0]), _p(pts),
             out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))) == 0
    return out


def unicycle_gate_flags(cost_blob, xy):
    """Bit j of row i: the unicycle obstacle cost's square gate of slot j
    holds point i (the plugin's unicycle_gate, as the GPU evaluates it)."""
    return _prefilter("oracle_unicycle_gate_flags", cost_blob, xy)


def bicycle_candidates(cost_blob, xy):
    """Bit j of row i: the bicycle pursuit cost's zone grid lists zone j for
    point i (the plugin's frame, band sets and band mapping)."""
    return _prefilter("oracle_bicycle_candidates", cost_blob, xy)


def philox4x32(ctr, key, library=False):
    """library: False (this oracle's own generator), True (the library
    header's compiled for the host) or "split" (the header's split form)."""
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    out = np.empty(4, np.uint32)
    fn = {False: lib().oracle_philox4x32, True: lib().oracle_library_philox4x32,
          "split": lib().oracle_library_philox4x32_split}[library]
    fn(_p(c, _U), _p(k, _U), _p(out, _U))
    return [int(v) for v in out]


def philox2x32(ctr, key):
    c = np.asarray(ctr, np.uint32)
    out = np.empty(2, np.uint32)
    lib().oracle_philox2x32(_p(c, _U), key, _p(out, _U))
    return [int(v) for v in out]


def normals4(block, sample, step, tag, agent, seed, fp64=False):
    if fp64:
        out = np.empty(4, np.float64)
        lib().oracle_normals4_64(block, sample, step, tag, agent, seed, _p(out, _D))
    else:
        out = np.empty(4, np.float32)
        lib().oracle_normals4(block, sample, step, tag, agent, seed, _p(out))
    return out


def mathf(fn, x):
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty_like(x)
    lib().oracle_mathf(fn, _p(x), _p(y), x.size)
    return y


def math2f(fn, x, y):
    """mppi_maxf (0), mppi_minf (1), mppi_clampf(x, -1/2, y) (2), host build."""
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    out = np.empty_like(x)
    lib().oracle_math2f(fn, _p(x), _p(y), _p(out), x.size)
    return out


def fmodf(x, y):
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    out = np.empty_like(x)
    lib().oracle_fmodf(_p(x), _p(y), _p(out), x.size)
    return out
