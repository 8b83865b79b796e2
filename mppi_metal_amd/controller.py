"""Python host binding of the HIP MPPI backend (one C-ABI context).

`HipMPPI` wraps one `mppi_hip_ctx`: it owns a HIP stream on one device and a
linked plugin.  Its methods are the C ABI's (include/mppi_hip.h), with numpy
arrays in place of pointers; every call raises `MppiHipError` with the
library's message on failure.  The C++ `mppi_metal::MPPIDriver` facade is the
drop-in for the reference's driver; this class is what tests and the
benchmark use from Python.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

from . import _native

PLUGIN_DIR = Path(__file__).resolve().parent / "plugins"


class MppiHipError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[{_native.STATUS_NAMES.get(code, code)}] {message}")
        self.code = code
        self.message = message


def _fptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _f32(a, shape=None):
    if a is None:
        return None
    out = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    if shape is not None and out.size != int(np.prod(shape)):
        raise ValueError(f"expected {int(np.prod(shape))} values, got {out.size}")
    return out


def _blob(b):
    """bytes / numpy array / None -> (c_void_p, nbytes, keepalive)."""
    if b is None:
        return None, 0, None
    if isinstance(b, np.ndarray):
        arr = np.ascontiguousarray(b)
        return ctypes.c_void_p(arr.ctypes.data), arr.nbytes, arr
    buf = ctypes.create_string_buffer(bytes(b), len(b))
    return ctypes.cast(buf, ctypes.c_void_p), len(b), buf


class HipMPPI:
    """One MPPI controller context on one GPU."""

    def __init__(self, state_dim, control_dim, horizon, sample_count, lam, sigma, u_min, u_max,
                 covariance_full=False, device=-1, numeric_mode=0):
        """numeric_mode: _native.FP32 (0), FP16 (1) or MIXED (2)."""
        self.lib = _native.load()
        self.D, self.C, self.H, self.K = int(state_dim), int(control_dim), int(horizon), int(sample_count)
        self.lam = float(lam)
        n_sigma = self.C * self.C if covariance_full else self.C
        self._sigma = _f32(sigma, (n_sigma,))
        self._umin = _f32(u_min, (self.C,))
        self._umax = _f32(u_max, (self.C,))
        cfg = _native.MppiHipConfig(self.D, self.C, self.H, self.K, self.lam,
                                    _native.COV_FULL if covariance_full else _native.COV_DIAGONAL,
                                    _fptr(self._sigma), _fptr(self._umin), _fptr(self._umax),
                                    int(numeric_mode), int(device))
        ctx = ctypes.c_void_p()
        rc = self.lib.mppi_hip_create(ctypes.byref(cfg), ctypes.byref(ctx))
        if rc != 0:
            raise MppiHipError(rc, self.lib.mppi_hip_last_error(None).decode())
        self.ctx = ctx
        self.model_bytes = 0
        self.cost_bytes = 0

    # -- lifecycle ---------------------------------------------------------
    def close(self):
        if getattr(self, "ctx", None):
            self.lib.mppi_hip_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise MppiHipError(rc, self.lib.mppi_hip_last_error(self.ctx).decode())

    # -- plugin and parameters --------------------------------------------
    def link(self, plugin, dynamics="mppi_dynamics", stage_cost="", terminal_cost="",
             model_bytes=0, cost_bytes=0, seed=0):
        """Compile and link a plugin file (a name like "unicycle.hip" is looked
        up among the bundled plugins)."""
        p = Path(plugin)
        if not p.exists() and (PLUGIN_DIR / plugin).exists():
            p = PLUGIN_DIR / plugin
        self._check(self.lib.mppi_hip_link_plugin(
            self.ctx, str(p).encode(), dynamics.encode(), (stage_cost or "").encode(),
            (terminal_cost or "").encode(), model_bytes, cost_bytes, int(seed) & (2**64 - 1)))
        self.model_bytes, self.cost_bytes = model_bytes, cost_bytes

    def link_source(self, source, dynamics="mppi_dynamics", stage_cost="", terminal_cost="",
                    model_bytes=0, cost_bytes=0, seed=0, name="<source>"):
        self._check(self.lib.mppi_hip_link_plugin_source(
            self.ctx, source.encode(), name.encode(), dynamics.encode(),
            (stage_cost or "").encode(), (terminal_cost or "").encode(), model_bytes, cost_bytes,
            int(seed) & (2**64 - 1)))
        self.model_bytes, self.cost_bytes = model_bytes, cost_bytes

    def bind_params(self, model=None, cost=None):
        mp, mn, mk = _blob(model)
        cp, cn, ck = _blob(cost)
        self._check(self.lib.mppi_hip_bind_params(self.ctx, mp, mn, cp, cn))

    # -- MPPI --------------------------------------------------------------
    def step(self, step_index, x0, u_nominal, model_override=None, cost_override=None):
        """One blocking iteration; returns (u_next [H, C], diag [3])."""
        x0 = _f32(x0, (self.D,))
        un = _f32(u_nominal, (self.H * self.C,))
        mp, mn, mk = _blob(model_override)
        cp, cn, ck = _blob(cost_override)
        u_out = np.empty(self.H * self.C, dtype=np.float32)
        diag = np.empty(3, dtype=np.float32)
        self._check(self.lib.mppi_hip_step(self.ctx, int(step_index), _fptr(x0), _fptr(un), mp, mn,
                                           cp, cn, _fptr(u_out), _fptr(diag)))
        return u_out.reshape(self.H, self.C), diag

    def step_noalloc(self, step_index, x0_dev, u_nominal_dev, u_out_dev, diag_dev=0, stream=0):
        """mppi_hip_step_noalloc: device addresses (integers) of the inputs
        and outputs; stream 0 waits on the context's stream, any other value
        is a hipStream_t to enqueue on."""
        self._check(self.lib.mppi_hip_step_noalloc(self.ctx, int(step_index), x0_dev or None, u_nominal_dev or None,
                                                   u_out_dev or None, diag_dev or None, stream or None))

    def simulate(self, num_steps, start_step, x0, u_nominal, model_override=None,
                 cost_override=None, outputs=("states", "controls", "costs")):
        x0 = _f32(x0, (self.D,))
        un = _f32(u_nominal, (self.H * self.C,))
        mp, mn, mk = _blob(model_override)
        cp, cn, ck = _blob(cost_override)
        T = int(num_steps)
        st = np.empty((T, self.D), np.float32) if "states" in outputs else None
        ct = np.empty((T, self.C), np.float32) if "controls" in outputs else None
        kt = np.empty(T, np.float32) if "costs" in outputs else None
        self._check(self.lib.mppi_hip_simulate(self.ctx, T, int(start_step), _fptr(x0), _fptr(un),
                                               mp, mn, cp, cn, _fptr(st), _fptr(ct), _fptr(kt)))
        return st, ct, kt

    def simulate_batch(self, num_agents, num_steps, start_step, x0_packed, u_nominal_packed=None,
                       outputs=("states", "controls", "costs", "terminal")):
        N, T = int(num_agents), int(num_steps)
        x0 = _f32(x0_packed, (N * self.D,))
        un = _f32(u_nominal_packed, (N * self.H * self.C,))
        st = np.empty((T, N, self.D), np.float32) if "states" in outputs else None
        ct = np.empty((T, N, self.C), np.float32) if "controls" in outputs else None
        kt = np.empty((T, N), np.float32) if "costs" in outputs else None
        tt = np.empty((N, self.D), np.float32) if "terminal" in outputs else None
        self._check(self.lib.mppi_hip_simulate_batch(self.ctx, N, T, int(start_step), _fptr(x0),
                                                     _fptr(un), _fptr(st), _fptr(ct), _fptr(kt),
                                                     _fptr(tt)))
        return st, ct, kt, tt

    def read_costs(self, agent=0, count=None):
        """Sample costs of the last iteration of `agent` (count = its K)."""
        n = self.K if count is None else int(count)
        out = np.empty(n, np.float32)
        self._check(self.lib.mppi_hip_read_costs(self.ctx, agent, _fptr(out), n))
        return out

    def read_weights(self, agent=0, count=None):
        n = self.K if count is None else int(count)
        out = np.empty(n, np.float32)
        self._check(self.lib.mppi_hip_read_weights(self.ctx, agent, _fptr(out), n))
        return out

    def set_step_path(self, path):
        """_native.PATH_AUTO, PATH_TWO_KERNEL or PATH_FUSED."""
        self._check(self.lib.mppi_hip_set_step_path(self.ctx, int(path)))

    def set_wait_limit(self, polls=0, force_timeout=False):
        """Bound of the one-pass rollout's wait for its noise (0: default);
        force_timeout makes every wait give up (tests the error path)."""
        self._check(self.lib.mppi_hip_set_wait_limit(self.ctx, int(polls), 1 if force_timeout else 0))

    def set_reduce_samples_per_thread(self, spt):
        self._check(self.lib.mppi_hip_set_reduce_samples_per_thread(self.ctx, int(spt)))

    # -- learned-dynamics surrogate (mppi_hip_surrogate_* in include/mppi_hip.h) --
    def surrogate_train(self, state_lo, state_hi, state_input=None, num_transitions=1 << 20,
                        batch=4096, steps=2000, learning_rate=2e-3, seed=1):
        """Generate transitions with the plugin dynamics and train the MLP on
        the GPU; returns the per-step batch MSE (normalised units)."""
        lo, hi = _f32(state_lo, (self.D,)), _f32(state_hi, (self.D,))
        mode = None if state_input is None else np.ascontiguousarray(state_input, np.uint32)
        cfg = _native.MppiHipSurrogateConfig(
            _fptr(lo), _fptr(hi),
            None if mode is None else mode.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
            int(num_transitions), int(batch), int(steps), float(learning_rate), int(seed))
        hist = np.empty(int(steps), np.float32)
        self._check(self.lib.mppi_hip_surrogate_train(self.ctx, ctypes.byref(cfg), _fptr(hist)))
        return hist

    def surrogate_evaluate(self, num_samples=65536, seed=7):
        """(rmse, relative rmse) per state component of one predicted step."""
        rmse, rel = np.empty(self.D, np.float32), np.empty(self.D, np.float32)
        self._check(self.lib.mppi_hip_surrogate_evaluate(self.ctx, int(num_samples), int(seed),
                                                         _fptr(rmse), _fptr(rel)))
        return rmse, rel

    def surrogate_predict(self, states, controls):
        s = _f32(states).reshape(-1, self.D)
        u = _f32(controls, (s.shape[0] * self.C,))
        out = np.empty_like(s)
        self._check(self.lib.mppi_hip_surrogate_predict(self.ctx, s.shape[0], _fptr(s), _fptr(u), _fptr(out)))
        return out

    def surrogate_use(self, enable=True):
        self._check(self.lib.mppi_hip_surrogate_use(self.ctx, 1 if enable else 0))

    def surrogate_export(self):
        n = ctypes.c_size_t()
        self._check(self.lib.mppi_hip_surrogate_export(self.ctx, None, 0, ctypes.byref(n), None))
        params = np.empty(n.value, np.float32)
        norm = np.empty(_native.SURROGATE_NORM_FLOATS, np.float32)
        self._check(self.lib.mppi_hip_surrogate_export(self.ctx, _fptr(params), params.size,
                                                       ctypes.byref(n), _fptr(norm)))
        return params, norm

    def surrogate_import(self, params, norm):
        p = _f32(params)
        nb = _f32(norm, (_native.SURROGATE_NORM_FLOATS,))
        self._check(self.lib.mppi_hip_surrogate_import(self.ctx, _fptr(p), p.size, _fptr(nb)))

    def surrogate_gradients(self, x, targets):
        """Summed gradients of the batch MSE (no update) and the loss."""
        xs = _f32(x)
        n = xs.shape[0]
        ts = _f32(targets, (n * self.D,))
        n_params = ctypes.c_size_t()
        self._check(self.lib.mppi_hip_surrogate_export(self.ctx, None, 0, ctypes.byref(n_params), None))
        g = np.empty(n_params.value, np.float32)
        loss = np.empty(1, np.float32)
        self._check(self.lib.mppi_hip_surrogate_gradients(self.ctx, n, _fptr(xs), _fptr(ts), _fptr(g),
                                                          _fptr(loss)))
        return g, float(loss[0])

    def surrogate_activation(self, z):
        """The network's hidden activation T of float32 pre-activations,
        evaluated on the device (binary16 results as float32)."""
        zs = _f32(z).ravel()
        out = np.empty_like(zs)
        self._check(self.lib.mppi_hip_surrogate_activation(self.ctx, zs.size, _fptr(zs), _fptr(out)))
        return out

    def profile(self, enable=True):
        self._check(self.lib.mppi_hip_profile_enable(self.ctx, 1 if enable else 0))

    def profile_read(self):
        r, d, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_uint64()
        self._check(self.lib.mppi_hip_profile_read(self.ctx, ctypes.byref(r), ctypes.byref(d),
                                                   ctypes.byref(n)))
        return r.value, d.value, n.value


class HipBatch:
    """Uneven multi-agent batch bound to a HipMPPI context
    (mppi_hip_batch_* in include/mppi_hip.h)."""

    def __init__(self, ctx: HipMPPI, sample_counts, horizons):
        self.ctx = ctx
        self.lib = ctx.lib
        self.K = np.ascontiguousarray(sample_counts, np.uint32)
        self.H = np.ascontiguousarray(horizons, np.uint32)
        self.N = int(self.K.size)
        if self.H.size != self.N:
            raise ValueError("sample_counts and horizons differ in length")
        self.seq = (self.H.astype(np.int64) * ctx.C)
        self.seq_off = np.concatenate([[0], np.cumsum(self.seq)[:-1]]).astype(np.int64)
        self.seq_total = int(self.seq.sum())
        u32p = ctypes.POINTER(ctypes.c_uint32)
        b = ctypes.c_void_p()
        ctx._check(self.lib.mppi_hip_batch_create(ctx.ctx, self.N, self.K.ctypes.data_as(u32p),
                                                  self.H.ctypes.data_as(u32p), ctypes.byref(b)))
        self.b = b

    def close(self):
        if getattr(self, "b", None):
            self.lib.mppi_hip_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, model=None, cost=None, model_stride=0, cost_stride=0):
        mp, mn, mk = _blob(model)
        cp, cn, ck = _blob(cost)
        self.ctx._check(self.lib.mppi_hip_batch_set_params(self.b, mp, model_stride, cp, cost_stride))

    def set_tuning(self, block=256, persistent=0, wgs_per_cu=4, reduce_spt=0, order=0, launches=2):
        t = _native.MppiHipBatchTuning(block, persistent, wgs_per_cu, reduce_spt, order, launches)
        self.ctx._check(self.lib.mppi_hip_batch_set_tuning(self.b, ctypes.byref(t)))

    def get_tuning(self):
        t = _native.MppiHipBatchTuning()
        self.ctx._check(self.lib.mppi_hip_batch_get_tuning(self.b, ctypes.byref(t)))
        return dict(block=t.block, persistent=t.persistent, wgs_per_cu=t.wgs_per_cu,
                    reduce_spt=t.reduce_spt, order=t.order, launches=t.launches)

    def autotune(self, trial_steps=3):
        ms = ctypes.c_double()
        self.ctx._check(self.lib.mppi_hip_batch_autotune(self.b, trial_steps, ctypes.byref(ms)))
        return ms.value

    def reset(self, start_step, x0_packed, u_nominal_packed=None):
        x0 = _f32(x0_packed, (self.N * self.ctx.D,))
        un = _f32(u_nominal_packed, (self.seq_total,))
        self.ctx._check(self.lib.mppi_hip_batch_reset(self.b, int(start_step), _fptr(x0), _fptr(un)))

    def run(self, num_steps, closed_loop=True, history=False):
        T, N, D, C = int(num_steps), self.N, self.ctx.D, self.ctx.C
        st = np.empty((T, N, D), np.float32) if history else None
        ct = np.empty((T, N, C), np.float32) if history else None
        kt = np.empty((T, N), np.float32) if history else None
        self.ctx._check(self.lib.mppi_hip_batch_run(self.b, T, 1 if closed_loop else 0, _fptr(st),
                                                    _fptr(ct), _fptr(kt)))
        return st, ct, kt

    def read(self):
        x = np.empty((self.N, self.ctx.D), np.float32)
        u = np.empty(self.seq_total, np.float32)
        d = np.empty((self.N, 3), np.float32)
        self.ctx._check(self.lib.mppi_hip_batch_read(self.b, _fptr(x), _fptr(u), _fptr(d)))
        return x, u, d

    def u_of(self, u_packed, a):
        o = int(self.seq_off[a])
        return u_packed[o:o + int(self.seq[a])].reshape(int(self.H[a]), self.ctx.C)


class HipServer:
    """Serving: agents admitted against an HBM budget, solved in batches
    (mppi_hip_server_* in include/mppi_hip.h)."""

    def __init__(self, ctx: HipMPPI, max_agents: int, hbm_budget_bytes: int):
        self.ctx, self.lib = ctx, ctx.lib
        self.s = ctypes.c_void_p()
        ctx._check(self.lib.mppi_hip_server_create(ctx.ctx, int(max_agents), int(hbm_budget_bytes),
                                                   ctypes.byref(self.s)))

    def close(self):
        if self.s:
            self.lib.mppi_hip_server_destroy(self.s)
            self.s = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def workspace_bytes(self, K, H):
        n = ctypes.c_size_t()
        self.ctx._check(self.lib.mppi_hip_server_workspace_bytes(self.s, int(K), int(H), ctypes.byref(n)))
        return n.value

    def admit(self, K, H, x0, u_nominal=None, cost=None):
        x = _f32(x0, (self.ctx.D,))
        un = _f32(u_nominal, (int(H) * self.ctx.C,)) if u_nominal is not None else None
        cp, cn, keep = _blob(cost)
        aid = ctypes.c_uint32()
        self.ctx._check(self.lib.mppi_hip_server_admit(self.s, int(K), int(H), _fptr(x), _fptr(un), cp, cn,
                                                       ctypes.byref(aid)))
        return aid.value

    def release(self, agent_id):
        self.ctx._check(self.lib.mppi_hip_server_release(self.s, int(agent_id)))

    def solve(self, agent_ids):
        ids = np.ascontiguousarray(agent_ids, np.uint32)
        n = ids.size
        u0 = np.empty((n, self.ctx.C), np.float32)
        diag = np.empty((n, 3), np.float32)
        self.ctx._check(self.lib.mppi_hip_server_solve(self.s, n, ids.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                                       _fptr(u0), _fptr(diag)))
        return u0, diag

    def read(self, agent_id, H):
        x = np.empty(self.ctx.D, np.float32)
        u = np.empty((int(H), self.ctx.C), np.float32)
        self.ctx._check(self.lib.mppi_hip_server_read(self.s, int(agent_id), _fptr(x), _fptr(u)))
        return x, u

    def usage(self):
        used, budget, agents = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_uint32()
        self.ctx._check(self.lib.mppi_hip_server_usage(self.s, ctypes.byref(used), ctypes.byref(budget),
                                                       ctypes.byref(agents)))
        return used.value, budget.value, agents.value

    def replay(self, arrival_s, K, H, x0, cost_blobs, u_nominal=None, solves=1, max_batch=0):
        """Real-time serving loop over a request trace; returns (latency_s [n],
        stats dict)."""
        n = len(arrival_s)
        x0 = np.ascontiguousarray(x0, np.float32).reshape(n, self.ctx.D)
        reqs = (_native.MppiHipRequest * n)()
        keep = []
        for i in range(n):
            r = reqs[i]
            r.arrival_s = float(arrival_s[i])
            r.K, r.H = int(K[i]), int(H[i])
            r.solves = int(solves)
            r.x0 = _fptr(x0[i])
            if u_nominal is not None:
                un = np.ascontiguousarray(u_nominal[i], np.float32)
                keep.append(un)
                r.u_nominal = _fptr(un)
            cp, _, k = _blob(cost_blobs[i] if cost_blobs is not None else None)
            keep.append(k)
            r.cost_params = cp
        lat = np.empty(n, np.float64)
        st = _native.MppiHipReplayStats()
        self.ctx._check(self.lib.mppi_hip_server_replay(self.s, n, reqs, int(max_batch),
                                                        lat.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                        ctypes.byref(st)))
        stats = {k: getattr(st, k) for k, _ in _native.MppiHipReplayStats._fields_}
        return lat, stats
