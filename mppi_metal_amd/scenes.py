"""Synthetic benchmark scenes and parameter blobs for the bundled plugins.

Scenes are drawn on the host, once, from Philox4x32-10 with the run's seed and
the scene tag in counter word 3 (block i, sample 0, step 0, tag 3 << 28), so
the CPU oracle and the GPU see byte-identical parameter blobs.

Unicycle workloads (the model of the reference's
examples/nonholonomic_vehicle/nonholonomic_vehicle.metal:24-91, weights from
examples/nonholonomic_vehicle/main.cpp:71-80):
  dt = 0.05, v_max = 2, sigma = [0.5, 0.3], lambda = 1, bounds +-[2, 1],
  goal ~ U[-10, 10]^2, n obstacles with centres ~ U[-10, 10]^2 (rejected when
  within r + 0.5 of x0 or the goal) and radius ~ U[0.5, 1.5];
  stage cost 10 |p - g|^2 + 0.1 |u|^2 + 1000 sum_i max(0, r_i - |p - c_i|),
  terminal cost 100 |p - g|^2.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np

TAG_SCENE = 3
UNICYCLE_MAX_OBSTACLES = 16
UNICYCLE_MODEL_BYTES = 8
UNICYCLE_COST_BYTES = 8 * 4 + UNICYCLE_MAX_OBSTACLES * 12


def philox4x32_10(ctr, key):
    """Philox4x32-10 on Python integers (Random123 constants)."""
    c = [int(v) & 0xFFFFFFFF for v in ctr]
    k0, k1 = int(key[0]) & 0xFFFFFFFF, int(key[1]) & 0xFFFFFFFF
    for _ in range(10):
        p0 = 0xD2511F53 * c[0]
        p1 = 0xCD9E8D57 * c[2]
        c = [((p1 >> 32) ^ c[1] ^ k0) & 0xFFFFFFFF, p1 & 0xFFFFFFFF,
             ((p0 >> 32) ^ c[3] ^ k1) & 0xFFFFFFFF, p0 & 0xFFFFFFFF]
        k0 = (k0 + 0x9E3779B9) & 0xFFFFFFFF
        k1 = (k1 + 0xBB67AE85) & 0xFFFFFFFF
    return c


def u01(word):
    """(0, 1) uniform of a 32-bit word, as the library converts it."""
    return float(np.float32(((int(word) >> 8) | 1) * 2.0 ** -24))


class SceneStream:
    """Sequential uniforms of the scene stream of one seed."""

    def __init__(self, seed: int, agent: int = 0):
        self.key = (seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)
        self.agent = agent
        self.block = 0
        self.buf = []

    def uniform(self, lo=0.0, hi=1.0):
        if not self.buf:
            w = philox4x32_10((self.block, 0, 0, (TAG_SCENE << 28) | self.agent), self.key)
            self.buf = [u01(x) for x in w]
            self.block += 1
        u = self.buf.pop(0)
        return lo + (hi - lo) * u


@dataclass
class UnicycleScene:
    goal: tuple
    obstacles: list                 # [(cx, cy, r)]
    v_max: float = 2.0
    dt: float = 0.05
    pos_weight: float = 10.0
    control_weight: float = 0.1
    obstacle_weight: float = 1000.0
    goal_radius: float = 0.2
    x0: tuple = (0.0, 0.0, 0.0, 0.0)
    sigma: tuple = (0.5, 0.3)
    u_min: tuple = (-2.0, -1.0)
    u_max: tuple = (2.0, 1.0)
    lam: float = 1.0
    extra: dict = field(default_factory=dict)

    def model_blob(self) -> bytes:
        return struct.pack("<2f", self.v_max, self.dt)

    def cost_blob(self) -> bytes:
        obs = list(self.obstacles)[:UNICYCLE_MAX_OBSTACLES]
        head = struct.pack("<7fI", self.goal[0], self.goal[1], self.goal_radius, self.pos_weight,
                           0.0, self.control_weight, self.obstacle_weight, len(obs))
        flat = []
        for i in range(UNICYCLE_MAX_OBSTACLES):
            flat.extend(obs[i] if i < len(obs) else (0.0, 0.0, 0.0))
        return head + struct.pack(f"<{3 * UNICYCLE_MAX_OBSTACLES}f", *flat)


def unicycle_scene(n_obstacles: int, seed: int = 42) -> UnicycleScene:
    s = SceneStream(seed)
    goal = (s.uniform(-10, 10), s.uniform(-10, 10))
    x0 = (0.0, 0.0)
    obstacles = []
    tries = 0
    while len(obstacles) < n_obstacles and tries < 10000:
        tries += 1
        cx, cy, r = s.uniform(-10, 10), s.uniform(-10, 10), s.uniform(0.5, 1.5)
        if np.hypot(cx - x0[0], cy - x0[1]) < r + 0.5 or np.hypot(cx - goal[0], cy - goal[1]) < r + 0.5:
            continue
        obstacles.append((float(np.float32(cx)), float(np.float32(cy)), float(np.float32(r))))
    goal = (float(np.float32(goal[0])), float(np.float32(goal[1])))
    return UnicycleScene(goal=goal, obstacles=obstacles)


# Benchmark workloads of the unicycle: (name, n_obstacles, H, K, closed-loop steps)
UNICYCLE_C1 = dict(name="unicycle_k1024_h32_obs8", n_obstacles=8, H=32, K=1024, steps=200)
UNICYCLE_C2 = dict(name="unicycle_k65536_h64_obs16", n_obstacles=16, H=64, K=65536, steps=500)


# ---------------------------------------------------------------------------
# Dynamic-bicycle pursuit (batched multi-agent workload)
# ---------------------------------------------------------------------------
BICYCLE_MAX_ZONES = 32
BICYCLE_MODEL_BYTES = 48
BICYCLE_COST_BYTES = 48 + BICYCLE_MAX_ZONES * 12


@dataclass
class BicycleParams:
    mass: float = 1000.0
    yaw_inertia: float = 1600.0
    lf: float = 1.0
    lr: float = 1.5
    cf: float = 4.0e4
    cr: float = 4.0e4
    drag: float = 0.1
    v_min: float = 2.0
    v_max: float = 20.0
    dt: float = 0.05
    substeps: int = 2

    def blob(self) -> bytes:
        return struct.pack("<10fIf", self.mass, self.yaw_inertia, self.lf, self.lr, self.cf,
                           self.cr, self.drag, self.v_min, self.v_max, self.dt, self.substeps, 0.0)


def bicycle_cost_blob(target, target_vel, zones, w_pos=1.0, w_control=0.1, w_zone=200.0,
                      w_terminal=10.0, dt=0.05) -> bytes:
    zs = list(zones)[:BICYCLE_MAX_ZONES]
    head = struct.pack("<9fI2f", target[0], target[1], target_vel[0], target_vel[1], w_pos,
                       w_control, w_zone, w_terminal, dt, len(zs), 0.0, 0.0)
    flat = []
    for i in range(BICYCLE_MAX_ZONES):
        flat.extend(zs[i] if i < len(zs) else (0.0, 0.0, 0.0))
    return head + struct.pack(f"<{3 * BICYCLE_MAX_ZONES}f", *flat)


@dataclass
class PursuitScene:
    model: BicycleParams
    zones: list
    x0: np.ndarray            # [N, 6]
    targets: np.ndarray       # [N, 4] (x, y, vx, vy)
    K: np.ndarray             # [N]
    H: np.ndarray             # [N]
    cost_blobs: list          # N blobs of BICYCLE_COST_BYTES
    sigma: tuple = (0.2, 1.5)
    u_min: tuple = (-0.5, -4.0)
    u_max: tuple = (0.5, 4.0)
    lam: float = 10.0

    def cost_blob_packed(self) -> bytes:
        return b"".join(self.cost_blobs)


def bicycle_pursuit_scene(n_agents=256, seed=42, k_log2=(10, 14), h_range=(16, 96),
                          n_zones=32) -> PursuitScene:
    """Agents start in [-40, 40]^2 outside the threat zones (centres in
    [-50, 50]^2, radius U[2, 8]) with vx = 5 m/s; each pursues its own target
    moving at a constant velocity of at most 3 m/s.  K_a = round(2^U[10, 14]),
    H_a uniform in [16, 96]."""
    s = SceneStream(seed)
    zones = []
    for _ in range(n_zones):
        zones.append((float(np.float32(s.uniform(-50, 50))), float(np.float32(s.uniform(-50, 50))),
                      float(np.float32(s.uniform(2, 8)))))
    x0 = np.zeros((n_agents, 6), np.float32)
    targets = np.zeros((n_agents, 4), np.float32)
    K = np.zeros(n_agents, np.uint32)
    H = np.zeros(n_agents, np.uint32)
    for a in range(n_agents):
        while True:
            px, py = s.uniform(-40, 40), s.uniform(-40, 40)
            if all(np.hypot(px - zx, py - zy) > zr + 1.0 for zx, zy, zr in zones):
                break
        x0[a] = (px, py, s.uniform(-np.pi, np.pi), 5.0, 0.0, 0.0)
        ang = s.uniform(-np.pi, np.pi)
        spd = s.uniform(0.0, 3.0)
        targets[a] = (s.uniform(-40, 40), s.uniform(-40, 40), spd * np.cos(ang), spd * np.sin(ang))
        K[a] = int(round(2.0 ** s.uniform(*k_log2)))
        H[a] = int(h_range[0] + min(h_range[1] - h_range[0], int(s.uniform(0, 1) * (h_range[1] - h_range[0] + 1))))
    blobs = [bicycle_cost_blob(targets[a, :2], targets[a, 2:], zones) for a in range(n_agents)]
    return PursuitScene(BicycleParams(), zones, x0, targets, K, H, blobs)


# Operating envelope of the learned bicycle surrogate's training data: states
# uniform over this box (positions are not network inputs: the dynamics is
# translation invariant; the heading is periodic), controls over the bounds.
BICYCLE_SURROGATE_ENVELOPE = dict(state_lo=[-50.0, -50.0, -np.pi, 0.0, -4.0, -4.0],
                                  state_hi=[50.0, 50.0, np.pi, 20.0, 4.0, 4.0],
                                  state_input=[2, 2, 1, 0, 0, 0])
# The training of the learned-surrogate workload (K = 2^18, H = 64): 4M
# transitions, 40000 Adam steps of 4096 (about forty passes over the data),
# peak step size 1e-3.
BICYCLE_SURROGATE_TRAINING = dict(num_transitions=4 << 20, batch=4096, steps=40000, learning_rate=1e-3,
                                  seed=7)


def pursuit_realised_cost(scene: PursuitScene, agent: int, next_states, controls, steps,
                          w_pos=1.0, w_control=0.1, w_zone=200.0, dt=0.05) -> np.ndarray:
    """The pursuit stage cost (plugins/bicycle.hip, bicycle_pursuit_cost) of
    what a closed loop did: for each closed-loop step s, the state the plant
    reached with the applied control, at time tau = (s + 1) dt; FP64."""
    x = np.asarray(next_states, np.float64).reshape(-1, 6)
    u = np.asarray(controls, np.float64).reshape(-1, 2)
    tau = (np.asarray(steps, np.float64) + 1.0) * dt
    tx, ty, tvx, tvy = (float(v) for v in scene.targets[agent])
    ex = x[:, 0] - (tx + tvx * tau)
    ey = x[:, 1] - (ty + tvy * tau)
    cost = w_pos * (ex * ex + ey * ey) + w_control * (u[:, 0] ** 2 + u[:, 1] ** 2)
    for cx, cy, r in scene.zones:
        d = np.hypot(x[:, 0] - cx, x[:, 1] - cy)
        cost += w_zone * np.where(d < r, r - d, 0.0)
    return cost


# ---------------------------------------------------------------------------
# Quadrotor serving workload (plugins/quadrotor.hip)
# ---------------------------------------------------------------------------
@dataclass
class QuadParams:
    gravity: float = 9.81
    drag: float = 0.1
    tau: float = 0.05
    dt: float = 0.02
    substeps: int = 2

    def blob(self) -> bytes:
        return struct.pack("<4fI3f", self.gravity, self.drag, self.tau, self.dt, self.substeps,
                           0.0, 0.0, 0.0)


QUAD_MODEL_BYTES = 32
QUAD_COST_BYTES = 48
QUAD_U_MIN = (0.0, -3.0, -3.0, -3.0)      # thrust [m/s^2], body rates [rad/s]
QUAD_U_MAX = (20.0, 3.0, 3.0, 3.0)
QUAD_LAMBDA = 10.0


def quad_cost_blob(goal, w_pos=4.0, w_vel=0.5, w_att=1.0, w_thrust=0.1, w_rate=0.05,
                   w_terminal=20.0) -> bytes:
    return struct.pack("<3f6f3f", goal[0], goal[1], goal[2], w_pos, w_vel, w_att, w_thrust, w_rate,
                       w_terminal, 0.0, 0.0, 0.0)


def quadrotor_covariance(seed: int = 42, C: int = 4, eig=(0.1, 1.0)) -> np.ndarray:
    """Random SPD control covariance Q diag(lambda) Q^T: lambda ~ U[eig],
    Q orthogonal from the QR factorisation of a Gaussian matrix (scene
    stream of `seed`, agent word 1)."""
    s = SceneStream(seed, agent=1)
    g = np.array([[np.sqrt(-2 * np.log(s.uniform())) * np.cos(2 * np.pi * s.uniform()) for _ in range(C)]
                  for _ in range(C)])
    q, r = np.linalg.qr(g)
    q = q * np.sign(np.diag(r))
    lam = np.array([s.uniform(*eig) for _ in range(C)])
    cov = (q * lam) @ q.T
    return ((cov + cov.T) / 2).astype(np.float32)


@dataclass
class QuadRequests:
    arrival_s: np.ndarray     # [n] arrival times (seconds from the start)
    K: np.ndarray             # [n]
    H: np.ndarray             # [n]
    x0: np.ndarray            # [n, 12]
    goals: np.ndarray         # [n, 3]

    def cost_blob(self, i: int) -> bytes:
        return quad_cost_blob(self.goals[i])


def quadrotor_requests(n: int = 1024, seed: int = 42, rate_hz: float = 2000.0, k_log2=(12, 15),
                       h_range=(30, 80)) -> QuadRequests:
    """n agent requests arriving as a Poisson process of `rate_hz`: K =
    round(2^U[12, 15]) (4096..32768), H uniform in [30, 80], start near hover
    (position U[-5, 5]^3, velocity U[-1, 1]^3, attitude U[-0.2, 0.2]^3, zero
    rates), goal U[-5, 5]^3."""
    s = SceneStream(seed, agent=2)
    t, arr = 0.0, np.zeros(n)
    K, H = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
    x0 = np.zeros((n, 12), np.float32)
    goals = np.zeros((n, 3), np.float32)
    for i in range(n):
        t += -np.log(s.uniform()) / rate_hz
        arr[i] = t
        K[i] = int(round(2.0 ** s.uniform(*k_log2)))
        H[i] = int(h_range[0] + min(h_range[1] - h_range[0], int(s.uniform() * (h_range[1] - h_range[0] + 1))))
        x0[i, 0:3] = [s.uniform(-5, 5) for _ in range(3)]
        x0[i, 3:6] = [s.uniform(-1, 1) for _ in range(3)]
        x0[i, 6:9] = [s.uniform(-0.2, 0.2) for _ in range(3)]
        goals[i] = [s.uniform(-5, 5) for _ in range(3)]
    return QuadRequests(arr, K, H, x0, goals)
