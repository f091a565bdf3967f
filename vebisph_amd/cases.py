"""Seeded synthetic inputs for the benchmark workloads and the open-domain (free-surface) cases.

The reference reads its initial state from IPART/INDAT/IBOUND files written by its case generator
(source/gen/datclass.F90).  This module builds the same kind of state in memory:

* boundary patches are the closed polygon b_node(1..nb) with b_edge(ib) = b_node(ib+1) - b_node(ib)
  (datclass.F90:730-744), walls type 1, periodic type 2 with b_periodic_parent;
* wall particles are laid along every wall patch at spacing dx, with the single extra node particle
  where a wall is followed by a non-wall patch, exactly as make_boundary_particles does
  (datclass.F90:746-815), and circle particles at ceiling(2*pi*R/dx) points (datclass.F90:761-773);
* fluid particles sit on the lattice ((i+1/2)dx, (j+1/2)dx), x outer / y inner loop order (the
  generator's ordering, datclass.F90:261-274), each coordinate jittered by U(-0.01dx, +0.01dx);
* the random stream is SplitMix64(seed) -> uniform double in [0,1) = (z >> 11) * 2^-53, consumed in
  fluid-particle index order: x-jitter, y-jitter, then (thermal cases) the temperature noise.

Particle order is the reference's: circle particles, then wall particles (indices 1..nbS), then fluid.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

NPLINK = 75          # per-cell / per-particle neighbour cap, common_parameter.F90:5


def splitmix64_uniform(seed: int, count: int) -> np.ndarray:
    """count uniforms in [0,1) from SplitMix64 started at `seed` (Steele, Lea & Flood 2014)."""
    with np.errstate(over="ignore"):
        k = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


@dataclasses.dataclass
class Case:
    name: str
    # run parameters (INDAT names, input.F90:18-37)
    h: float
    dx: float
    dv: float
    Ra: float
    Pr: float
    El: float
    beta: float
    body_force: tuple
    dt_coef_cfl: float = 0.4
    dt_coef_visc: float = 0.2
    ve_nonlinearity: float = 0.0
    i_open_domain: int = 0
    i_shift_coeff: int = 0
    # the reference's compile-time fluid model: 2 viscoelastic (linear PTT / Oldroyd-B), 1 Newtonian
    const_mod: int = 2
    steps: int = 1
    # boundaries (IBOUND, input.F90:51-81)
    b_node: np.ndarray = None        # (nb, 2)
    b_type: np.ndarray = None        # (nb,)
    b_vel: np.ndarray = None
    b_thermal: np.ndarray = None
    b_periodic_parent: np.ndarray = None
    c_centre: np.ndarray = None      # (nc, 2)
    c_radius: np.ndarray = None
    c_omega: np.ndarray = None
    c_vel: np.ndarray = None         # (nc, 2)
    c_thermal: np.ndarray = None
    # particles 0..npfb-1 (IPART, input.F90:139-144)
    nbS: int = 0
    x: np.ndarray = None
    y: np.ndarray = None
    u: np.ndarray = None
    v: np.ndarray = None
    theta: np.ndarray = None

    @property
    def npfb(self) -> int:
        return int(self.x.shape[0])

    @property
    def nb_patches(self) -> int:
        return int(self.b_node.shape[0])

    @property
    def nb_circles(self) -> int:
        return 0 if self.c_radius is None else int(self.c_radius.shape[0])

    @property
    def b_edge(self) -> np.ndarray:
        return np.roll(self.b_node, -1, axis=0) - self.b_node

    @property
    def Wi(self) -> float:
        return self.El * math.sqrt(self.Ra / self.Pr)

    @property
    def Re(self) -> float:
        return math.sqrt(self.Ra / self.Pr)

    def params_vector(self) -> np.ndarray:
        return np.array([self.h, self.dx, self.dv, self.Ra, self.Pr, self.El, self.beta,
                         self.body_force[0], self.body_force[1], self.dt_coef_cfl,
                         self.dt_coef_visc, self.ve_nonlinearity], dtype=np.float64)


def _boundary_particles(b_node, b_type, b_vel, b_thermal, dx, c_centre=None, c_radius=None,
                        c_omega=None, c_vel=None, c_thermal=None):
    """make_boundary_particles (datclass.F90:746-815), same floating-point expression order."""
    nb = b_node.shape[0]
    b_edge = np.roll(b_node, -1, axis=0) - b_node
    xs, ys, us, vs, ts = [], [], [], [], []
    if c_radius is not None:
        for ib in range(c_radius.shape[0]):
            R = abs(float(c_radius[ib]))
            tmp = 2.0 * R * math.pi
            nround = math.ceil(tmp / dx)
            for iround in range(1, nround + 1):
                tmp2 = float(iround) * 2.0 * math.pi / float(nround)
                xs.append(c_centre[ib, 0] + R * math.cos(tmp2))
                ys.append(c_centre[ib, 1] + R * math.sin(tmp2))
                us.append(c_vel[ib, 0] + c_omega[ib] * R * math.sin(tmp2))
                vs.append(c_vel[ib, 1] - c_omega[ib] * R * math.cos(tmp2))
                ts.append(c_thermal[ib])
    for ib in range(nb):
        ibm1 = (ib - 1) % nb
        if abs(b_type[ib]) == 1:
            m_be = math.sqrt(b_edge[ib, 0] * b_edge[ib, 0] + b_edge[ib, 1] * b_edge[ib, 1])
            x, y = b_node[ib, 0], b_node[ib, 1]
            tmp = 0.0
            while tmp < 1.0 - 1.0e-10:
                xs.append(x); ys.append(y)
                uu = b_vel[ib] * b_edge[ib, 0] / m_be
                vv = b_vel[ib] * b_edge[ib, 1] / m_be
                if abs(b_type[ibm1]) == 1 and b_vel[ibm1] != 0.0 and tmp == 0.0:
                    mp = math.sqrt(b_edge[ibm1, 0] ** 2 + b_edge[ibm1, 1] ** 2)
                    uu = b_vel[ibm1] * b_edge[ibm1, 0] / mp
                    vv = b_vel[ibm1] * b_edge[ibm1, 1] / mp
                us.append(uu); vs.append(vv); ts.append(b_thermal[ib])
                tmp = tmp + dx / m_be
                x = b_node[ib, 0] + tmp * b_edge[ib, 0]
                y = b_node[ib, 1] + tmp * b_edge[ib, 1]
        elif abs(b_type[ibm1]) == 1:
            xs.append(b_node[ib, 0]); ys.append(b_node[ib, 1])
            m_be = math.sqrt(b_edge[ibm1, 0] ** 2 + b_edge[ibm1, 1] ** 2)
            us.append(b_vel[ibm1] * b_edge[ibm1, 0] / m_be)
            vs.append(b_vel[ibm1] * b_edge[ibm1, 1] / m_be)
            ts.append(b_thermal[ibm1])
    f = lambda a: np.asarray(a, dtype=np.float64)
    return f(xs), f(ys), f(us), f(vs), f(ts)


def channel_case(name, L, H, nx, ny, seed, *, Ra, Pr, El, beta, body_force=(0.0, 0.0),
                 theta_bottom=0.0, theta_top=0.0, theta_linear=False, theta_noise=0.0,
                 circle=None, steps=1, jitter=0.01, h_over_dx=1.3) -> Case:
    """Periodic-in-x channel [0,L]x[0,H] with walls at y=0 and y=H (patch types 1,2,1,2)."""
    dx = L / nx
    if abs(H / dx - ny) > 1e-9:
        raise ValueError("lattice must tile the channel: H/dx must equal ny")
    h = h_over_dx * dx
    b_node = np.array([[0.0, 0.0], [L, 0.0], [L, H], [0.0, H]])
    b_type = np.array([1, 2, 1, 2], dtype=np.int32)
    b_vel = np.zeros(4)
    b_thermal = np.array([theta_bottom, 0.0, theta_top, 0.0])
    b_pp = np.array([3, 4, 1, 2], dtype=np.int32)
    cc = cr = co = cv = ct = None
    if circle is not None:
        (cx, cy), R = circle
        cc = np.array([[cx, cy]]); cr = np.array([R]); co = np.zeros(1)
        cv = np.zeros((1, 2)); ct = np.zeros(1)
    bx, by, bu, bv, bt = _boundary_particles(b_node, b_type, b_vel, b_thermal, dx, cc, cr, co, cv, ct)
    ix, iy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")   # x outer, y inner
    fx = (ix.ravel() + 0.5) * dx
    fy = (iy.ravel() + 0.5) * dx
    if circle is not None:
        keep = np.hypot(fx - cc[0, 0], fy - cc[0, 1]) >= cr[0] + 0.25 * dx
        fx, fy = fx[keep], fy[keep]
    nf = fx.shape[0]
    per = 3 if theta_noise else 2
    rnd = splitmix64_uniform(seed, per * nf).reshape(nf, per)
    fx = fx + (2.0 * rnd[:, 0] - 1.0) * (jitter * dx)
    fy = fy + (2.0 * rnd[:, 1] - 1.0) * (jitter * dx)
    if theta_linear:
        ft = theta_bottom + (theta_top - theta_bottom) * fy / H
    else:
        ft = np.zeros(nf)
    if theta_noise:
        ft = ft + (2.0 * rnd[:, 2] - 1.0) * theta_noise
    c = Case(name=name, h=h, dx=dx, dv=dx * dx, Ra=Ra, Pr=Pr, El=El, beta=beta,
             body_force=tuple(body_force), steps=steps, b_node=b_node, b_type=b_type,
             b_vel=b_vel, b_thermal=b_thermal, b_periodic_parent=b_pp, c_centre=cc,
             c_radius=cr, c_omega=co, c_vel=cv, c_thermal=ct, nbS=bx.shape[0],
             x=np.concatenate([bx, fx]), y=np.concatenate([by, fy]),
             u=np.concatenate([bu, np.zeros(nf)]), v=np.concatenate([bv, np.zeros(nf)]),
             theta=np.concatenate([bt, ft]))
    return c


def _f32(v: float) -> float:
    """A REAL(4) literal of the generator widened to double (datclass.F90 mixes them into FP64)."""
    return float(np.float32(v))


def free_surface_convection(nx=150, L=1.0, H=1.0, depth=0.5, seed=6, Ra=2.0e4, Pr=1.0, El=0.01,
                            beta=0.5, eps_ptt=0.1, steps=1, jitter=0.01, name="fs_convection_11k") -> Case:
    """Free-surface Rayleigh-Benard flow (the generator's case 6, datclass.F90:433-488): a layer of
    depth `depth` on a hot wall (theta = 1) in [0,L]x[0,H], periodic in x (patch types 1,2,0,2, the
    top patch open), gravity (0,-1), open domain with the velocity-scaled shifting coefficient
    (i_shift_coeff 1).  theta = 1 below y = 0.2 depth + 0.1 depth sin(2 pi x / L), else 0 (the
    generator's REAL(4) constants)."""
    dx = L / nx
    h = 1.3 * dx
    b_node = np.array([[0.0, 0.0], [L, 0.0], [L, H], [0.0, H]])
    b_type = np.array([1, 2, 0, 2], dtype=np.int32)
    b_vel = np.zeros(4)
    b_thermal = np.array([1.0, 0.0, -1.0, 0.0])
    b_pp = np.array([3, 4, 1, 2], dtype=np.int32)
    bx, by, bu, bv, bt = _boundary_particles(b_node, b_type, b_vel, b_thermal, dx)
    ny = int(math.ceil(depth / dx - 0.5))
    while (ny - 0.5) * dx >= depth:
        ny -= 1
    ix, iy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")   # x outer, y inner
    fx = (ix.ravel() + 0.5) * dx
    fy = (iy.ravel() + 0.5) * dx
    nf = fx.shape[0]
    rnd = splitmix64_uniform(seed, 2 * nf).reshape(nf, 2)
    fx = fx + (2.0 * rnd[:, 0] - 1.0) * (jitter * dx)
    fy = fy + (2.0 * rnd[:, 1] - 1.0) * (jitter * dx)
    edge = _f32(0.2) * depth + _f32(0.1) * depth * np.sin(_f32(2.0) * math.pi * fx / L)
    ft = np.where(fy <= edge, 1.0, 0.0)
    return Case(name=name, h=h, dx=dx, dv=dx * dx, Ra=Ra, Pr=Pr, El=El, beta=beta,
                body_force=(0.0, -1.0), ve_nonlinearity=eps_ptt, i_open_domain=1, i_shift_coeff=1,
                steps=steps, b_node=b_node, b_type=b_type, b_vel=b_vel, b_thermal=b_thermal,
                b_periodic_parent=b_pp, nbS=bx.shape[0], x=np.concatenate([bx, fx]),
                y=np.concatenate([by, fy]), u=np.concatenate([bu, np.zeros(nf)]),
                v=np.concatenate([bv, np.zeros(nf)]), theta=np.concatenate([bt, np.zeros(nf) + ft]))


def dam_break(ncol=40, h0=0.25, seed=7, Ra=1.0e4, Pr=1.0, El=0.1, beta=0.5, eps_ptt=0.1,
              shift_coeff=0, steps=1, jitter=0.01, name="dam_break_3k") -> Case:
    """A viscoelastic dam break: the geometry of the generator's case 1 (datclass.F90:56-121) -- a
    column [0,h0]x[0,2 h0] collapsing under gravity (0,-1) in the closed box [0,1]^2 (four walls),
    dx = h0/ncol -- with a polymeric phase (El > 0, beta < 1; the generator's case is Newtonian)."""
    dx = h0 / ncol
    h = 1.3 * dx
    b_node = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])
    b_type = np.array([1, 1, 1, 1], dtype=np.int32)
    b_vel = np.zeros(4)
    b_thermal = np.zeros(4)
    b_pp = np.zeros(4, dtype=np.int32)
    bx, by, bu, bv, bt = _boundary_particles(b_node, b_type, b_vel, b_thermal, dx)
    ix, iy = np.meshgrid(np.arange(ncol), np.arange(2 * ncol), indexing="ij")
    fx = (ix.ravel() + 0.5) * dx
    fy = (iy.ravel() + 0.5) * dx
    nf = fx.shape[0]
    rnd = splitmix64_uniform(seed, 2 * nf).reshape(nf, 2)
    fx = fx + (2.0 * rnd[:, 0] - 1.0) * (jitter * dx)
    fy = fy + (2.0 * rnd[:, 1] - 1.0) * (jitter * dx)
    return Case(name=name, h=h, dx=dx, dv=dx * dx, Ra=Ra, Pr=Pr, El=El, beta=beta,
                body_force=(0.0, -1.0), ve_nonlinearity=eps_ptt, i_open_domain=1,
                i_shift_coeff=shift_coeff, steps=steps, b_node=b_node, b_type=b_type, b_vel=b_vel,
                b_thermal=b_thermal, b_periodic_parent=b_pp, nbS=bx.shape[0],
                x=np.concatenate([bx, fx]), y=np.concatenate([by, fy]),
                u=np.concatenate([bu, np.zeros(nf)]), v=np.concatenate([bv, np.zeros(nf)]),
                theta=np.concatenate([bt, np.zeros(nf)]))


def poiseuille(nx=50, ny=50, L=1.0, H=1.0, seed=1, Wi=1.0, beta=0.1, Re=0.1, steps=1000, name="poiseuille_2500"):
    """Oldroyd-B start-up Poiseuille flow, body force (8,0), theta = 0 (config 1; config 4 shape)."""
    Ra = Re * Re          # Pr = 1, Re = sqrt(Ra/Pr)
    return channel_case(name, L, H, nx, ny, seed, Ra=Ra, Pr=1.0, El=Wi / Re, beta=beta,
                        body_force=(8.0, 0.0), steps=steps)


def convection(nx=400, ny=200, L=2.0, H=1.0, seed=2, Ra=1e4, Wi=0.5, beta=0.5, steps=2000, name="convection_80k"):
    """Boussinesq viscoelastic convection: hot bottom (+1), cold top (-1), linear T + 1e-3 noise."""
    return channel_case(name, L, H, nx, ny, seed, Ra=Ra, Pr=1.0, El=Wi / math.sqrt(Ra), beta=beta,
                        theta_bottom=1.0, theta_top=-1.0, theta_linear=True, theta_noise=1e-3,
                        steps=steps)


def cylinder(nx=1432, ny=716, seed=3, Wi=1.0, beta=0.59, Re=0.01, f=8.0, steps=500, name="cylinder_1M"):
    """Flow past a fixed cylinder (R=0.5 at (2,2)) in an 8x4 periodic channel, body force (f,0)."""
    return channel_case(name, 8.0, 4.0, nx, ny, seed, Ra=Re * Re, Pr=1.0, El=Wi / Re, beta=beta,
                        body_force=(f, 0.0), circle=((2.0, 2.0), 0.5), steps=steps)


CONFIGS = {
    "poiseuille_2500": lambda: poiseuille(),
    "convection_80k": lambda: convection(),
    "cylinder_1M": lambda: cylinder(),
    # the same flow at twice and four times the resolution (capacity and throughput at size)
    "cylinder_2M": lambda: cylinder(nx=2026, ny=1013, name="cylinder_2M"),
    "cylinder_4M": lambda: cylinder(nx=2864, ny=1432, name="cylinder_4M"),
    "cylinder_16M": lambda: cylinder(nx=5728, ny=2864, name="cylinder_16M"),
    "poiseuille_250k": lambda: poiseuille(nx=1000, ny=250, L=4.0, H=1.0, seed=4, steps=200,
                                          name="poiseuille_250k"),
    "convection_1M": lambda: convection(nx=1414, ny=707, seed=5, Ra=5e4, Wi=1.0, beta=0.5, steps=500,
                                        name="convection_1M"),
    "fs_convection_11k": lambda: free_surface_convection(),
    "fs_convection_1M": lambda: free_surface_convection(nx=1414, seed=8, name="fs_convection_1M"),
    "dam_break_3k": lambda: dam_break(),
}


def make(name: str) -> Case:
    return CONFIGS[name]()
