"""The reference's case generator: the eight cases of source/gen/datclass.F90, written as run
directories (INDAT, IBOUND, IPART) that the reference's program and `vb_drive run` read.

    python -m vebisph_amd.generator CASE DIR        (CASE 1..8, as datclass.F90 asks on stdin)

Each case builds the generator's geometry, parameters and particles with its floating-point
operations in the generator's order (lattices by repeated addition, the REAL(4) literals widened:
`_f32`), and the files carry the generator's layout: INDAT in its fixed formats (F18.10 and
friends, which round h, dx, dt_out and the physical groups to ten decimals -- the values a run then
uses), IBOUND and IPART list-directed.  Where the generator leaves a value unset (case 1: the fluid
temperature and b_thermal; case 2: no b_vel before make_boundary_particles zeroes it) the files
hold 0, which is what the generator writes here.

  1  dam break                 datclass.F90:45-121    open domain, Newtonian (El = 0)
  2  Taylor-Green vortices     :123-198               all periodic, Newtonian
  3  Poiseuille flow           :200-275               walls y, periodic x
  4  Rayleigh-Benard           :277-348               hot/cold walls
  5  flow past a cylinder      :350-430               Grilli et al. 2013, one circle
  6  free-surface R-B          :432-488               open domain, i_shift_coeff 1
  7  plane Couette             :490-546               moving walls
  8  free-surface channel      :548-610               open domain, i_shift_coeff 1
"""
from __future__ import annotations

import dataclasses
import math
import os
import sys

import numpy as np

from vebisph_amd import cases, refio

PI = 3.14159265358979323846          # common_parameter.F90:12


def _f32(v: float) -> float:
    return float(np.float32(v))


@dataclasses.dataclass
class GenRun:
    case: cases.Case
    tmax: float
    dt_out: float
    alpha_evss: float = 0.0


def _square(xl, yl, origin=(0.0, 0.0)):
    x0, y0 = origin
    return np.array([[x0, y0], [x0 + xl, y0], [x0 + xl, y0 + yl], [x0, y0 + yl]])


def _build(name, *, xl, yl, nodes, b_type, b_pp, b_vel, b_thermal, dx, h, fluid, params, circles=None,
           tmax, dt_out, open_domain=0, shift_coeff=0):
    """make_boundary_edge_vectors + make_boundary_particles (datclass.F90:730-815), then the fluid
    particles from `fluid(xb_min, xb_max, yb_min, yb_max)` -> lists x, y, u, v, theta."""
    b_node = np.asarray(nodes, dtype=np.float64)
    cc = cr = co = cv = ct = None
    if circles:
        cc = np.array([c[0] for c in circles], dtype=np.float64)
        cr = np.array([c[1] for c in circles], dtype=np.float64)
        co = np.array([c[2] for c in circles], dtype=np.float64)
        cv = np.array([c[3] for c in circles], dtype=np.float64)
        ct = np.array([c[4] for c in circles], dtype=np.float64)
    b_type = np.asarray(b_type, dtype=np.int32)
    b_vel = np.asarray(b_vel, dtype=np.float64)
    b_thermal = np.asarray(b_thermal, dtype=np.float64)
    bx, by, bu, bv, bt = cases._boundary_particles(b_node, b_type, b_vel, b_thermal, dx, cc, cr, co, cv, ct)
    xb_min, xb_max = b_node[:, 0].min(), b_node[:, 0].max()
    yb_min, yb_max = b_node[:, 1].min(), b_node[:, 1].max()
    fx, fy, fu, fv, ft = fluid(xb_min, xb_max, yb_min, yb_max)
    f = lambda a: np.asarray(a, dtype=np.float64)
    Ra, Pr, El, beta, eps, bf, cfl, visc = params
    c = cases.Case(name=name, h=h, dx=dx, dv=dx * dx, Ra=Ra, Pr=Pr, El=El, beta=beta, body_force=tuple(bf),
                   dt_coef_cfl=cfl, dt_coef_visc=visc, ve_nonlinearity=eps, i_open_domain=open_domain,
                   i_shift_coeff=shift_coeff, b_node=b_node, b_type=b_type, b_vel=b_vel, b_thermal=b_thermal,
                   b_periodic_parent=np.asarray(b_pp, dtype=np.int32), c_centre=cc, c_radius=cr, c_omega=co,
                   c_vel=cv, c_thermal=ct, nbS=bx.shape[0],
                   x=np.concatenate([bx, f(fx)]), y=np.concatenate([by, f(fy)]),
                   u=np.concatenate([bu, f(fu)]), v=np.concatenate([bv, f(fv)]),
                   theta=np.concatenate([bt, f(ft)]))
    return GenRun(case=c, tmax=tmax, dt_out=dt_out)


def case1():
    """2-D dam break (datclass.F90:45-121)."""
    xl = yl = 1.0; h0 = 0.25
    dx = h0 / 40.0; h = 1.3 * dx

    def fluid(x0, x1, y0, y1):
        X, Y = [], []
        x = x0 + dx * 0.5
        while x < x1 - 0.5 * dx:
            y = y0 + dx * 0.5
            while y < y1 - 0.5 * dx:
                if x <= h0 and y <= 2.0 * h0:
                    X.append(x); Y.append(y)
                y = y + dx
            x = x + dx
        n = len(X)
        return X, Y, [0.0] * n, [0.0] * n, [0.0] * n
    return _build("gen1_dam_break", xl=xl, yl=yl, nodes=_square(xl, yl), b_type=[1, 1, 1, 1],
                  b_pp=[3, 4, 1, 2], b_vel=[0.0] * 4, b_thermal=[0.0] * 4, dx=dx, h=h, fluid=fluid,
                  params=(1.0e4, 1.0, 0.0, 1.0, 100.0, (0.0, -1.0), 0.4, 0.2),
                  tmax=100.0, dt_out=_f32(0.01), open_domain=1)


def case2():
    """Taylor-Green vortices (datclass.F90:123-198): no walls, u = -cos(2 pi x) sin(2 pi y)."""
    xl = yl = 1.0
    dx = yl / 100.0; h = 1.3 * dx

    def fluid(x0, x1, y0, y1):
        X, Y = [], []
        y = y0 + dx / 2.0
        while y < y1:
            x = x0 + dx / 2.0
            while x < x1:
                X.append(x); Y.append(y)
                x = x + dx
            y = y + dx
        U = [-math.cos(2.0 * x * PI) * math.sin(2.0 * y * PI) for x, y in zip(X, Y)]
        V = [math.sin(2.0 * x * PI) * math.cos(2.0 * y * PI) for x, y in zip(X, Y)]
        return X, Y, U, V, [0.0] * len(X)
    return _build("gen2_taylor_green", xl=xl, yl=yl, nodes=_square(xl, yl), b_type=[2, 2, 2, 2],
                  b_pp=[3, 4, 1, 2], b_vel=[-0.0, 0.0, 0.0, 0.0], b_thermal=[1.0, 0.0, -1.0, 0.0], dx=dx,
                  h=h, fluid=fluid, params=(1.0e4, 1.0, 0.0, 1.0, 100.0, (0.0, 0.0), 0.4, 0.2),
                  tmax=100.0, dt_out=_f32(0.01))


def _channel_fluid(dx, theta_of):
    """the channel lattice of cases 3, 4, 8: x from xb_min + dx/2, y from yb_min + dx while
    y < yb_max - dx/2 (case 8: y < ymax_fluid)"""
    def fluid(x0, x1, y0, y1, ytop=None):
        X, Y, T = [], [], []
        x = x0 + 0.5 * dx
        while x < x1:
            y = y0 + dx
            while (y < y1 - 0.5 * dx) if ytop is None else (y < ytop):
                X.append(x); Y.append(y); T.append(theta_of(x, y))
                y = y + dx
            x = x + dx
        n = len(X)
        return X, Y, [0.0] * n, [0.0] * n, T
    return fluid


def case3():
    """Poiseuille flow (datclass.F90:200-275)."""
    xl = yl = 1.0
    dx = yl / 30.0; h = 1.3 * dx
    fl = _channel_fluid(dx, lambda x, y: 0.0 if y <= 0.5 * yl + 0.25 * yl * math.sin(2.0 * PI * x) else -0.0)
    return _build("gen3_poiseuille", xl=xl, yl=yl, nodes=_square(xl, yl), b_type=[1, 2, 1, 2],
                  b_pp=[3, 4, 1, 2], b_vel=[-0.0, 0.0, 0.0, 0.0], b_thermal=[0.0, 0.0, -0.0, 0.0], dx=dx, h=h,
                  fluid=fl, params=(1.0, 1.0, 1.0, 0.1, 0.0, (8.0, 0.0), 0.4, 0.2),
                  tmax=100.0, dt_out=_f32(0.1))


def case4():
    """Rayleigh-Benard flow (datclass.F90:277-348): theta = 1 below y = 1/2 + sin(2 pi x)/4, else -1."""
    xl = yl = 1.0
    dx = yl / 100.0; h = 1.3 * dx
    fl = _channel_fluid(dx, lambda x, y: 1.0 if y <= 0.5 * yl + 0.25 * yl * math.sin(2.0 * PI * x) else -1.0)
    return _build("gen4_rayleigh_benard", xl=xl, yl=yl, nodes=_square(xl, yl), b_type=[1, 2, 1, 2],
                  b_pp=[3, 4, 1, 2], b_vel=[-0.0, 0.0, 0.0, 0.0], b_thermal=[1.0, 0.0, -1.0, 0.0], dx=dx, h=h,
                  fluid=fl, params=(1.0e4, 1.0, 1.0e-4, 0.9, 0.1, (0.0, 0.0), 0.4, 0.2),
                  tmax=100.0, dt_out=_f32(0.1))


def case5():
    """Poiseuille flow past a cylinder (Grilli et al. 2013; datclass.F90:350-430)."""
    xl, yl, h0 = 3.0, 2.0, 0.5
    dx = h0 / 20.0; h = 1.3 * dx
    cen, rad = (0.0, 0.0), 0.25 * yl

    def fluid(x0, x1, y0, y1):
        X, Y = [], []
        x = x0 + 0.5 * dx
        while x < x1:
            y = y0 + dx
            while y < y1 - 0.5 * dx:
                dxc, dyc, rr = x - cen[0], y - cen[1], rad + 0.25 * dx
                if dxc * dxc + dyc * dyc > rr * rr:
                    X.append(x); Y.append(y)
                y = y + dx
            x = x + dx
        n = len(X)
        return X, Y, [0.0] * n, [0.0] * n, [0.0] * n
    nodes = [[-0.5 * xl, -0.5 * yl], [0.5 * xl, -0.5 * yl], [0.5 * xl, 0.5 * yl], [-0.5 * xl, 0.5 * yl]]
    return _build("gen5_cylinder", xl=xl, yl=yl, nodes=nodes, b_type=[1, 2, 1, 2], b_pp=[3, 4, 1, 2],
                  b_vel=[0.0] * 4, b_thermal=[0.0] * 4, dx=dx, h=h, fluid=fluid,
                  params=(10000.0, 1.0, 20.0, 0.1, 100.0, (0.0, 0.0), 0.4, 0.2),
                  circles=[(cen, rad, 0.0, (0.0, 0.0), 1.0)], tmax=100.0, dt_out=_f32(0.01))


def case6():
    """Free-surface Rayleigh-Benard flow (datclass.F90:432-488)."""
    xl = yl = 1.0; h0 = 0.5
    dx = yl / 150.0; h = 1.3 * dx

    def fluid(x0, x1, y0, y1):
        X, Y, T = [], [], []
        x = x0 + 0.5 * dx
        while x < x1:
            y = y0 + 0.5 * dx
            while y < h0:
                X.append(x); Y.append(y)
                edge = _f32(0.2) * h0 + _f32(0.1) * h0 * math.sin(2.0 * PI * x / xl)
                T.append(1.0 if y <= edge else 0.0)
                y = y + dx
            x = x + dx
        n = len(X)
        return X, Y, [0.0] * n, [0.0] * n, T
    return _build("gen6_fs_rayleigh_benard", xl=xl, yl=yl, nodes=_square(xl, yl), b_type=[1, 2, 0, 2],
                  b_pp=[3, 4, 1, 2], b_vel=[0.0] * 4, b_thermal=[1.0, 0.0, -1.0, 0.0], dx=dx, h=h, fluid=fluid,
                  params=(2.0e4, 1.0, 0.01, 0.5, 0.1, (0.0, -1.0), 0.4, 0.2),
                  tmax=1000.0, dt_out=_f32(0.1), open_domain=1, shift_coeff=1)


def case7():
    """Plane Couette flow (datclass.F90:490-546): walls moving at -1 (bottom, along +x) and -1 (top,
    along -x); the fluid lattice runs downwards from the top."""
    xl = yl = 1.0
    dx = yl / 30.0; h = 1.3 * dx

    def fluid(x0, x1, y0, y1):
        X, Y = [], []
        x = x0 + 0.5 * dx
        while x < x1:
            y = y1 - dx
            while y > y0 + 0.5 * dx:
                X.append(x); Y.append(y)
                y = y - dx
            x = x + dx
        n = len(X)
        return X, Y, [0.0] * n, [0.0] * n, [0.0] * n
    nodes = [[-0.5 * xl, -0.5 * yl], [0.5 * xl, -0.5 * yl], [0.5 * xl, 0.5 * yl], [-0.5 * xl, 0.5 * yl]]
    return _build("gen7_couette", xl=xl, yl=yl, nodes=nodes, b_type=[1, 2, 1, 2], b_pp=[3, 4, 1, 2],
                  b_vel=[-1.0, 0.0, -1.0, 0.0], b_thermal=[1.0, 0.0, -1.0, 0.0], dx=dx, h=h, fluid=fluid,
                  params=(1.0, 1.0, 1.0, 0.1, 100.0, (0.0, 0.0), 0.4, 0.2), tmax=100.0, dt_out=_f32(0.01))


def case8():
    """Free-surface periodic channel (datclass.F90:548-610): a layer of depth yl/2 on a hot wall,
    theta = 1 below y = yl/8 + sin(2 pi x) yl/16, else -1."""
    xl = yl = 2.0
    dx = yl / 75.0; h = 1.3 * dx
    base = _channel_fluid(dx, lambda x, y: 1.0 if y <= 0.125 * yl + 0.0625 * yl * math.sin(2.0 * PI * x) else -1.0)

    def fluid(x0, x1, y0, y1):
        return base(x0, x1, y0, y1, ytop=0.5 * yl)
    return _build("gen8_fs_channel", xl=xl, yl=yl, nodes=_square(xl, yl), b_type=[1, 2, 0, 2],
                  b_pp=[3, 4, 1, 2], b_vel=[-0.0, 0.0, 0.0, 0.0], b_thermal=[1.0, 0.0, -1.0, 0.0], dx=dx, h=h,
                  fluid=fluid, params=(1000.0, 10.0, 1.0, 0.1, 400.0, (0.0, 0.0), 0.4, 0.2),
                  tmax=100.0, dt_out=_f32(0.1), open_domain=1, shift_coeff=1)


CASES = {1: case1, 2: case2, 3: case3, 4: case4, 5: case5, 6: case6, 7: case7, 8: case8}


def _indat_text(run: GenRun) -> str:
    """INDAT in the generator's formats (datclass.F90:662-682)."""
    c = run.case
    F = lambda v: f"{v:18.10f}"
    lines = [" F" + " " + "        :i_RESTART",          # list-directed: a blank before each item
             f"{c.i_open_domain:5d}        :open domain (0-closed,1-open)",
             f"{F(run.tmax)}        :tmax - max time of simulation",
             f"{c.i_shift_coeff:5d}        :shift_coeff - how much to shift",
             f"{c.npfb:7d}        :np", f"{c.npfb:7d}        :npfb", f"{c.nbS:7d}        :nbS",
             f"{F(run.dt_out)} :dt_out - time interval between outputs",
             f"{F(c.h)} :h", f"{F(c.dx)} :dx", f"{F(0.0)} :dt",
             f"{F(c.body_force[0])}{c.body_force[1]:18.6f} :body_force",
             f"{F(c.Ra)} Rayleigh Number", f"{F(c.Pr)} Prandtl Number", f"{F(c.El)} Prandtl Number",
             f"{F(c.beta)} Viscosity ratio", f"{F(c.dt_coef_cfl)}{F(c.dt_coef_visc)} :dt_coefs cfl,visc",
             f"{F(c.ve_nonlinearity)}  :ve_nonlinearity parameter", f"{F(run.alpha_evss)}  :EVSS coefficient"]
    return "\n".join(lines) + "\n"


def write_run_dir(k: int, path: str) -> GenRun:
    """Generator case k (1..8) as a run directory: INDAT, IBOUND, IPART."""
    run = CASES[k]()
    refio.write_case_dir(path, run.case, tmax=run.tmax, dt_out=run.dt_out)     # IBOUND, IPART
    with open(os.path.join(path, "INDAT"), "w") as f:                         # the generator's formats
        f.write(_indat_text(run))
    return run


def main(argv):
    if len(argv) != 3 or argv[1] not in {str(k) for k in CASES}:
        print(__doc__.split("\n\n")[1])
        return 2
    run = write_run_dir(int(argv[1]), argv[2])
    print(f"{run.case.name}: np = {run.case.npfb} nbS = {run.case.nbS} dx {run.case.dx!r} h {run.case.h!r}")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
