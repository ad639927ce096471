"""Oracle (test infrastructure only): NumPy FP64 restatement of the host
rigid-body solver of ``iceflow3d/rigid.py:121-404``.

Only tests/, the smoke check and bench.py's CPU-baseline leg use it.  It is
written against the reference algorithm, independently of the product's
native C++ solver, so that the two can be compared on the reference's own
fixture states.
"""

from itertools import combinations, product

import numpy as np

CORNER_SIGNS = np.array(list(product((-1.0, 1.0), repeat=3)))


def rot(q):
    """rigid.py:14-22."""
    w, x, y, z = np.asarray(q, dtype=float) / np.linalg.norm(q)
    return np.array(
        [
            [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
            [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
            [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
        ]
    )


def qmul(a, b):
    """rigid.py:25-35 (Hamilton product)."""
    w, x, y, z = a
    v, i, j, k = b
    return np.array(
        [
            w * v - x * i - y * j - z * k,
            w * i + x * v + y * k - z * j,
            w * j - x * k + y * v + z * i,
            w * k + x * j - y * i + z * v,
        ]
    )


def qstep(q, omega, dt=1.0):
    """rigid.py:38-42."""
    angle = np.linalg.norm(omega) * dt
    scale = 0.5 * dt if angle < 1e-12 else np.sin(0.5 * angle) / np.linalg.norm(omega)
    out = qmul(np.r_[np.cos(0.5 * angle), scale * omega], q)
    return out / np.linalg.norm(out)


def sat(ca, ra, ha, cb, rb, hb):
    """rigid.py:45-62."""
    delta = ca - cb
    if np.any(abs(delta) > abs(ra) @ ha + abs(rb) @ hb + 1e-9):
        return None
    axes = [*ra.T, *rb.T] + [np.cross(a, b) for a in ra.T for b in rb.T]
    best_n, best_d = None, np.inf
    for axis in axes:
        length = np.linalg.norm(axis)
        if length < 1e-9:
            continue
        n = axis / length
        d = np.dot(ha, abs(ra.T @ n)) + np.dot(hb, abs(rb.T @ n)) - abs(delta @ n)
        if d < -1e-9:
            return None
        if d < best_d:
            best_n, best_d = (n if delta @ n >= 0 else -n), max(0.0, d)
    return best_n, best_d


def support_point(center, r, half, direction):
    """rigid.py:65-69."""
    local = r.T @ direction
    signs = np.where(abs(local) < 1e-9, 0.0, np.sign(local))
    return center + r @ (half * signs)


def impulse(va, wa, ma, ia, ra, n, vb=None, wb=None, mb=None, ib=None, rb=None):
    """rigid.py:72-89."""
    rel = va + np.cross(wa, ra)
    eff = 1 / ma + n @ np.cross(ia @ np.cross(ra, n), ra)
    if vb is not None:
        rel -= vb + np.cross(wb, rb)
        eff += 1 / mb + n @ np.cross(ib @ np.cross(rb, n), rb)
    closing = rel @ n
    if closing < 0:
        j = -closing / eff * n
        va += j / ma
        wa += ia @ np.cross(ra, j)
        if vb is not None:
            vb -= j / mb
            wb -= ib @ np.cross(rb, j)
    return max(0.0, -closing)


def manifold(v, w, m, inv_i, arms, n):
    """rigid.py:92-118."""
    ang = np.cross(arms, n)
    speed = (v + np.cross(w, arms)) @ n
    tol = 1e-12 * max(1.0, float(np.max(abs(speed))))
    if speed.min() >= -tol:
        return
    eff = 1 / m + ang @ inv_i @ ang.T
    for count in range(1, min(3, len(arms)) + 1):
        for subset in combinations(range(len(arms)), count):
            s = np.array(subset)
            imp = np.linalg.lstsq(eff[np.ix_(s, s)], -speed[s], rcond=1e-12)[0]
            if np.any(imp < 0):
                continue
            res = speed + eff[:, s] @ imp
            if np.any(res < -tol) or np.any(abs(res[s]) > tol):
                continue
            v += n * imp.sum() / m
            w += inv_i @ (ang[s].T @ imp)
            return
    raise RuntimeError("wall contact manifold has no feasible nonnegative impulse")


class OracleRigid:
    """rigid.py:121-404 with the simulation's device state in FP64."""

    def __init__(self, sim):
        self.s = sim
        bodies = sim.config.bodies
        self.origin = np.array([b.center_cells for b in bodies], dtype=float)
        self.q = np.array([b.quaternion_wxyz for b in bodies], dtype=float)
        self.velocity = np.array([b.velocity_lattice for b in bodies], dtype=float)
        self.omega = np.array([b.angular_velocity_lattice for b in bodies], dtype=float)
        self.com = np.zeros((len(bodies), 3))
        self.mass = np.zeros(len(bodies))
        self.inertia = np.zeros((len(bodies), 3, 3))
        self.lower = -0.5 * np.array([b.size_cells for b in bodies], dtype=float)
        self.upper = -self.lower
        self.upload()

    def upload(self):
        """rigid.py:136-151: the device receives the FP64 host state rounded
        once to FP32 (``state.astype(np.float32)``)."""
        s = self.s
        f32 = lambda v: np.asarray(v, dtype=np.float32).astype(float)  # noqa: E731
        for b in range(len(self.q)):
            r = rot(self.q[b])
            s.body_origin[b] = f32(self.origin[b])
            s.body_rotation[b] = f32(r)
            s.body_center[b] = f32(self.origin[b] + r @ self.com[b])
            s.body_velocity[b] = f32(self.velocity[b])
            s.body_angular_velocity[b] = f32(self.omega[b])

    def update_mass(self, *, upload=True):
        """rigid.py:153-176."""
        s = self.s
        s._reduce_mass()
        # rigid.py:156-158: downloaded through an FP32 buffer.
        red = s.reduced_state.to_numpy().astype(np.float32).astype(float)
        for b in range(len(red)):
            m, first = red[b, 0], red[b, 1:4]
            second = red[b, 4:13].reshape(3, 3)
            if m > 1e-15:
                com = first / m
                shift = rot(self.q[b]) @ (com - self.com[b])
                self.velocity[b] += np.cross(self.omega[b], shift)
                self.com[b] = com
                self.inertia[b] = second - m * (np.dot(com, com) * np.eye(3) - np.outer(com, com))
            else:
                self.velocity[b] = 0.0
                self.omega[b] = 0.0
        self.mass = red[:, 0].copy()
        self.lower = red[:, 13:16].copy()
        self.upper = red[:, 16:19].copy()
        if upload:
            self.upload()

    def advance(self):
        """rigid.py:178-211."""
        s = self.s
        # rigid.py:180-182: loads are downloaded through an FP32 buffer.
        forces = s.hydrodynamic_force.to_numpy().astype(np.float32).astype(float)
        torques = s.hydrodynamic_torque.to_numpy().astype(np.float32).astype(float)
        gravity = np.array(s.config.gravity) * s.scales.dt_s**2 / s.scales.dx_m
        for b in range(len(self.mass)):
            if self.mass[b] <= 1e-15:
                continue
            r = rot(self.q[b])
            inertia = r @ self.inertia[b] @ r.T
            center = self.origin[b] + r @ self.com[b]
            self.velocity[b] += forces[b] / self.mass[b] + gravity
            self.omega[b] += np.linalg.solve(
                inertia, torques[b] - np.cross(self.omega[b], inertia @ self.omega[b])
            )
            self.velocity[b] *= s.config.linear_damping
            self.omega[b] *= s.config.angular_damping
            travel = np.linalg.norm(self.velocity[b]) + np.linalg.norm(self.omega[b]) * np.linalg.norm(
                self.upper[b] - self.lower[b]
            ) / 2
            if not np.isfinite(travel) or travel > 0.5:
                raise RuntimeError(f"rigid displacement exceeds half a cell at step {s.steps + 1}, body {b}")
            center += self.velocity[b]
            self.q[b] = qstep(self.q[b], self.omega[b])
            self.origin[b] = center - rot(self.q[b]) @ self.com[b]
        self.resolve_contacts()
        self.upload()

    def _active(self):
        return [b for b in range(len(self.mass)) if self.mass[b] > 1e-15 and np.all(self.upper[b] > self.lower[b])]

    def _corners(self, b, r):
        return self.origin[b] + (self.lower[b] + (CORNER_SIGNS + 1) / 2 * (self.upper[b] - self.lower[b])) @ r.T

    def resolve_contacts(self):
        """rigid.py:213-347."""
        s = self.s
        low = np.full(3, s.config.boundary_cells, dtype=float)
        high = np.array(s.shape) - low
        active = self._active()
        rots, invs = None, None
        for _ in range(s.config.body_contact_iterations):
            worst, closing = 0.0, 0.0
            rots = [rot(q) for q in self.q]
            invs = {b: np.linalg.inv(rots[b] @ self.inertia[b] @ rots[b].T) for b in active}
            for b in active:
                r = rots[b]
                verts = self._corners(b, r)
                for axis, side in product(range(3), (-1, 1)):
                    ext = verts[:, axis].min() if side == 1 else verts[:, axis].max()
                    depth = low[axis] - ext if side == 1 else ext - high[axis]
                    if depth < -1e-8:
                        continue
                    n = np.eye(3)[axis] * side
                    shift = max(0.0, depth) * n
                    self.origin[b] += shift
                    verts += shift
                    worst = max(worst, depth)
                    target = verts[:, axis].min() if side == 1 else verts[:, axis].max()
                    for p in verts[abs(verts[:, axis] - target) < 1e-6]:
                        arm = p - (self.origin[b] + r @ self.com[b])
                        closing = max(
                            closing,
                            impulse(self.velocity[b], self.omega[b], self.mass[b], invs[b], arm, n),
                        )
            for i, a in enumerate(active):
                for b in active[i + 1 :]:
                    ra, rb = rots[a], rots[b]
                    ca = self.origin[a] + ra @ ((self.lower[a] + self.upper[a]) / 2)
                    cb = self.origin[b] + rb @ ((self.lower[b] + self.upper[b]) / 2)
                    ha = (self.upper[a] - self.lower[a]) / 2
                    hb = (self.upper[b] - self.lower[b]) / 2
                    hit = sat(ca, ra, ha, cb, rb, hb)
                    if hit is None:
                        continue
                    n, depth = hit
                    point = 0.5 * (support_point(ca, ra, ha, -n) + support_point(cb, rb, hb, n))
                    closing = max(
                        closing,
                        impulse(
                            self.velocity[a], self.omega[a], self.mass[a], invs[a],
                            point - self.origin[a] - ra @ self.com[a], n,
                            self.velocity[b], self.omega[b], self.mass[b], invs[b],
                            point - self.origin[b] - rb @ self.com[b],
                        ),
                    )
                    total = self.mass[a] + self.mass[b]
                    self.origin[a] += (depth + 1e-9) * n * self.mass[b] / total
                    self.origin[b] -= (depth + 1e-9) * n * self.mass[a] / total
                    worst = max(worst, depth)
            if worst < 1e-7 and closing < 1e-7:
                return
        worst, closing = self.contact_residuals()
        if worst < 1e-7 and closing < 1e-7:
            return
        if worst < 1e-7:
            for b in active:
                r = rots[b]
                center = self.origin[b] + r @ self.com[b]
                verts = self._corners(b, r)
                for axis, side in product(range(3), (-1, 1)):
                    ext = verts[:, axis].min() if side == 1 else verts[:, axis].max()
                    depth = low[axis] - ext if side == 1 else ext - high[axis]
                    if depth >= -1e-8:
                        pts = verts[abs(verts[:, axis] - ext) < 1e-6]
                        manifold(self.velocity[b], self.omega[b], self.mass[b], invs[b], pts - center,
                                 np.eye(3)[axis] * side)
            worst, closing = self.contact_residuals()
            if worst < 1e-7 and closing < 1e-7:
                return
        raise RuntimeError(
            f"contact projection did not converge at step {s.steps + 1}: "
            f"penetration={worst:.9g} cells, closing_speed={closing:.9g} LU"
        )

    def contact_residuals(self):
        """rigid.py:349-404."""
        s = self.s
        low = np.full(3, s.config.boundary_cells, dtype=float)
        high = np.array(s.shape) - low
        active = self._active()
        rots = [rot(q) for q in self.q]
        worst, closing = 0.0, 0.0
        for b in active:
            r = rots[b]
            verts = self._corners(b, r)
            center = self.origin[b] + r @ self.com[b]
            for axis, side in product(range(3), (-1, 1)):
                ext = verts[:, axis].min() if side == 1 else verts[:, axis].max()
                depth = low[axis] - ext if side == 1 else ext - high[axis]
                worst = max(worst, depth)
                if depth >= -1e-8:
                    pts = verts[abs(verts[:, axis] - ext) < 1e-6]
                    speeds = self.velocity[b] + np.cross(self.omega[b], pts - center)
                    closing = max(closing, float((-side * speeds[:, axis]).max()))
        for i, a in enumerate(active):
            for b in active[i + 1 :]:
                ra, rb = rots[a], rots[b]
                ca = self.origin[a] + ra @ ((self.lower[a] + self.upper[a]) / 2)
                cb = self.origin[b] + rb @ ((self.lower[b] + self.upper[b]) / 2)
                ha = (self.upper[a] - self.lower[a]) / 2
                hb = (self.upper[b] - self.lower[b]) / 2
                hit = sat(ca, ra, ha, cb, rb, hb)
                if hit is not None:
                    n, depth = hit
                    point = 0.5 * (support_point(ca, ra, ha, -n) + support_point(cb, rb, hb, n))
                    rel = (
                        self.velocity[a] + np.cross(self.omega[a], point - self.origin[a] - ra @ self.com[a])
                        - self.velocity[b] - np.cross(self.omega[b], point - self.origin[b] - rb @ self.com[b])
                    )
                    worst = max(worst, depth)
                    closing = max(closing, float(-rel @ n))
        return worst, closing
