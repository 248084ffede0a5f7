"""Case definitions: geometry, parameters and seeded initial fields.

A ``Case`` is everything ``ipl_create`` needs: the run configuration (the reference's Items /
items[] block plus the constants its drivers hard-code) and the initial macroscopic fields in
canonical cell order (x slowest, z fastest, as the reference's lnum loop,
LockExchange.cu:57-74).  Random fields come from SplitMix64 on the host in FP64, so every
consumer of a case sees identical bits.

The five workloads named by the project's baseline are built by ``lattice_taylor_green``,
``lock_exchange``, ``falling_sphere_flow``, ``solitary_wave`` and the serving mix in
``ip_cu_lbm_amd.serve``.
"""
from __future__ import annotations

import ctypes
import dataclasses
from typing import Dict, Optional

import numpy as np

from . import _abi

SEED_BASE = 20261203


def splitmix64(seed: int, n: int, stream: int = 0) -> np.ndarray:
    """n outputs of SplitMix64 (Steele, Lea, Flood 2014) started at seed + stream * 2^32."""
    gamma = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        s = np.uint64((seed + (stream << 32)) & 0xFFFFFFFFFFFFFFFF) + gamma * np.arange(1, n + 1, dtype=np.uint64)
        z = s
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, n: int, lo: float, hi: float, stream: int = 0) -> np.ndarray:
    """n doubles uniform in [lo, hi): the top 53 bits of SplitMix64."""
    u = (splitmix64(seed, n, stream) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    return lo + (hi - lo) * u


def axis_from_widths(widths: np.ndarray):
    widths = np.asarray(widths, dtype=np.float64)
    pos = np.cumsum(widths) - 0.5 * widths
    return pos, widths


def subdivided_axis(divs, coarse: float):
    """Piecewise refinement of the drivers (divx / divy, SphereFall.cu:80-128): coarse column i of
    width `coarse` is split into divs[i] equal cells, centred at (i + (2 d + 1) / (2 divs[i])) coarse.
    Returns (pos, widths)."""
    pos, wid = [], []
    for i, d in enumerate(divs):
        for k in range(int(d)):
            pos.append((i + (2.0 * k + 1.0) / (2.0 * d)) * coarse)
            wid.append(coarse / d)
    return np.array(pos), np.array(wid)


def sphere_fall_divs(n: int, coarse: float, inner=(0.03, 0.095), outer=(0.02, 0.105)):
    """The falling-sphere driver's refinement pattern (SphereFall.cu:80-104): columns whose centre
    lies in `inner` split in 4, in `outer` in 2, the rest kept."""
    divs = []
    for i in range(n):
        x = (i + 0.5) * coarse
        if outer[0] < x < outer[1]:
            divs.append(4 if inner[0] < x < inner[1] else 2)
        else:
            divs.append(1)
    return divs


def stretched_widths(n: int, ratio: float, dz: float = 1.0) -> np.ndarray:
    """Widths growing geometrically by `ratio` per cell away from the axis centre (finest at
    the centre): width_i = dz * ratio**j, j = cells between i and the centre pair."""
    i = np.arange(n, dtype=np.float64)
    j = np.floor(np.abs(i - (n - 1) / 2.0))
    return dz * ratio ** j


@dataclasses.dataclass
class Case:
    name: str
    lattice: int
    n: tuple  # (nx, ny, nz)
    pos: list  # 3 arrays
    width: list  # 3 arrays
    dz: float = 1.0
    dt: float = 1.0
    periodic: tuple = (0, 0, 0)
    wall_bc: int = _abi.BC_BOUNCE_BACK
    flux_axes: int = _abi.FLUX_REFERENCE
    stencil: int = 5
    collision: int = _abi.COLL_MRT
    s_rates: Optional[np.ndarray] = None
    tau: float = 0.6
    taus: float = 0.6
    rho_L: float = 1.0
    rho_H: float = 1.0
    beta_s: float = 0.0
    gravity: float = 0.0
    boussinesq: int = 1
    sigma: float = 0.0
    pf_thickness: float = 1.0
    scalar: int = _abi.SCALAR_NONE
    scalar_lattice: int = 0
    storage: int = _abi.F64
    path: int = _abi.PATH_AUTO
    f16_scale: float = 0.0
    check_finite: int = 0
    pf_tau: float = 0.58
    wall_function: int = 0
    wall_nu: float = 0.0
    les: int = 0
    mu_L: float = 0.0
    mu_H: float = 0.0
    line_axis: int = 0         # fused path, uniform grids: 1/2/3 = lines along x/y/z, 0 = automatic
    workgroup: int = 0         # fused path: work-items per workgroup (multiple of 64, <= 512), 0 = automatic
    solid: Optional[np.ndarray] = None
    init: Dict[str, np.ndarray] = dataclasses.field(default_factory=dict)
    steps: int = 100
    body: Optional["Body"] = None

    @property
    def n_cells(self) -> int:
        if self.solid is not None:
            return int(np.count_nonzero(self.solid == 0))
        return int(self.n[0] * self.n[1] * self.n[2])

    def coords(self):
        """Cell-centre coordinates of the calc cells in canonical order (X, Y, Z arrays)."""
        X, Y, Z = np.meshgrid(self.pos[0], self.pos[1], self.pos[2], indexing="ij")
        X, Y, Z = X.ravel(), Y.ravel(), Z.ravel()
        if self.solid is not None:
            keep = self.solid.ravel() == 0
            X, Y, Z = X[keep], Y[keep], Z[keep]
        return X, Y, Z

    def replace(self, **kw) -> "Case":
        return dataclasses.replace(self, **kw)

    # ---- ctypes marshalling (keeps the numpy buffers alive on the returned holder) ----------
    def to_c(self):
        keep = []

        def dp(a):
            a = np.ascontiguousarray(a, dtype=np.float64)
            keep.append(a)
            return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))

        cfg = _abi.IplConfig()
        cfg.lattice = self.lattice
        cfg.nx, cfg.ny, cfg.nz = self.n
        cfg.dz, cfg.dt = self.dz, self.dt
        for a in range(3):
            cfg.pos[a] = dp(self.pos[a])
            cfg.width[a] = dp(self.width[a])
        if self.solid is not None:
            s = np.ascontiguousarray(self.solid, dtype=np.uint8).ravel()
            keep.append(s)
            cfg.solid = s.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        for a in range(3):
            cfg.periodic[a] = int(self.periodic[a])
        cfg.wall_bc = self.wall_bc
        cfg.flux_axes = self.flux_axes
        cfg.stencil = self.stencil
        cfg.collision = self.collision
        rates = np.ones(27) if self.s_rates is None else np.asarray(self.s_rates, dtype=np.float64)
        for i in range(min(27, len(rates))):
            cfg.s_rates[i] = float(rates[i])
        cfg.tau, cfg.taus = self.tau, self.taus
        cfg.rho_L, cfg.rho_H = self.rho_L, self.rho_H
        cfg.beta_s, cfg.gravity = self.beta_s, self.gravity
        cfg.boussinesq = self.boussinesq
        cfg.sigma, cfg.pf_thickness = self.sigma, self.pf_thickness
        cfg.scalar, cfg.scalar_lattice = self.scalar, self.scalar_lattice
        cfg.storage, cfg.path = self.storage, self.path
        cfg.f16_scale = self.f16_scale
        cfg.check_finite = self.check_finite
        cfg.pf_tau = self.pf_tau
        cfg.wall_function = self.wall_function
        cfg.wall_nu = self.wall_nu
        cfg.les = self.les
        cfg.line_axis = self.line_axis
        cfg.workgroup = self.workgroup
        cfg.mu_L = self.mu_L
        cfg.mu_H = self.mu_H
        ini = _abi.IplInit()
        for key in ("p", "u", "v", "w", "S", "phi", "tau", "taus"):
            if key in self.init and self.init[key] is not None:
                arr = np.asarray(self.init[key], dtype=np.float64)
                if arr.shape != (self.n_cells,):
                    raise ValueError(f"init field {key} has shape {arr.shape}, expected ({self.n_cells},)")
                setattr(ini, key, dp(arr))
        return cfg, ini, keep


@dataclasses.dataclass
class Body:
    """A rigid body carried by Lagrangian markers (the drivers' marker set-up,
    SphereFall3D/SphereFall.cu:197-232): body-frame marker offsets and unit normals, spacing,
    mass properties and the initial state."""
    markers: np.ndarray           # (M, 3) offsets from the centre, body frame
    normals: np.ndarray           # (M, 3) outward unit normals, body frame
    spacing: float                # dIBM
    mass: float
    density: float
    inertia: tuple
    pos: tuple
    vel: tuple = (0.0, 0.0, 0.0)
    omega: tuple = (0.0, 0.0, 0.0)
    quat: tuple = (1.0, 0.0, 0.0, 0.0)
    gravity: tuple = (0.0, 0.0, 0.0)
    rho_f: float = 1.0
    method: int = _abi.IB_STRESS
    kernel: int = _abi.IB_KERNEL_REFERENCE
    iterations: int = 1
    moving: int = 1
    spm: int = 0                  # smoothed-profile forcing of the interior (SPM_ellipse3D)
    spm_a: float = 0.0
    spm_b: float = 0.0

    def to_c(self):
        keep = []

        def dp(a):
            a = np.ascontiguousarray(a, dtype=np.float64).ravel()
            keep.append(a)
            return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))

        c = _abi.IplIbConfig()
        c.method, c.kernel, c.iterations, c.moving = self.method, self.kernel, self.iterations, self.moving
        c.spm, c.spm_a, c.spm_b = self.spm, self.spm_a, self.spm_b
        c.n_markers = len(self.markers)
        c.marker_pos = dp(self.markers)
        c.marker_normal = dp(self.normals)
        c.spacing, c.rho_f, c.density, c.mass = self.spacing, self.rho_f, self.density, self.mass
        for a in range(3):
            c.inertia[a] = float(self.inertia[a])
            c.gravity[a] = float(self.gravity[a])
            c.pos[a] = float(self.pos[a])
            c.vel[a] = float(self.vel[a])
            c.omega[a] = float(self.omega[a])
        for i in range(4):
            c.quat[i] = float(self.quat[i])
        return c, keep


def icosphere(n: int) -> np.ndarray:
    """Unit-sphere points of an icosahedron whose faces are subdivided n times per edge and
    projected to the sphere (10 n^2 + 2 points): the marker sets of the reference's sphere drivers
    (make_3DSphere.py, sphere_points_n*.csv).  Marker spacing: the icosahedron edge / n,
    4 R / sqrt(10 + 2 sqrt 5) / n (SphereFall.cu:277)."""
    t = (1.0 + 5 ** 0.5) / 2.0
    V = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    V /= np.linalg.norm(V[0])
    F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    pts = {}
    for a, b, c in F:
        A, B, C = V[a], V[b], V[c]
        for i in range(n + 1):
            for j in range(n + 1 - i):
                p = A + (B - A) * (i / n) + (C - A) * (j / n)
                p = p / np.linalg.norm(p)
                pts.setdefault(tuple(np.round(p, 12)), p)
    out = np.array(list(pts.values()))
    return out[np.lexsort((out[:, 2], out[:, 1], out[:, 0]))]


def fibonacci_sphere(m: int) -> np.ndarray:
    """m unit-sphere points on a Fibonacci lattice: z_i = 1 - (2 i + 1) / m, azimuth i times the
    golden angle.  Mean spacing sqrt(4 pi / m) per unit radius."""
    i = np.arange(m, dtype=np.float64)
    z = 1.0 - (2.0 * i + 1.0) / m
    r = np.sqrt(1.0 - z * z)
    phi = i * np.pi * (3.0 - 5 ** 0.5)
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def sphere_body(radius: float, centre, markers: np.ndarray, spacing: float, density_ratio: float, rho_f: float = 1.0,
                gravity=(0.0, 0.0, 0.0), **kw) -> Body:
    """A solid sphere: markers on the surface (unit points times the radius), normals the unit
    points, mass and moment of inertia 2/5 m R^2 (SphereFall.cu:219-224)."""
    m = density_ratio * rho_f * 4.0 / 3.0 * np.pi * radius ** 3
    inertia = 2.0 / 5.0 * m * radius ** 2
    return Body(markers=markers * radius, normals=markers.copy(), spacing=spacing, mass=m, density=density_ratio * rho_f,
                inertia=(inertia, inertia, inertia), pos=tuple(float(c) for c in centre), gravity=tuple(gravity),
                rho_f=rho_f, **kw)


def falling_sphere(n=(192, 192, 384), diameter: float = 24.0, markers: int = 1800, density_ratio: float = 1.5,
                   tau: float = 0.644, gravity: float = 1.0e-4, iterations: int = 3, kernel: int = _abi.IB_KERNEL_PESKIN4,
                   method: int = _abi.IB_DIRECT, storage: int = _abi.F32, path: int = _abi.PATH_AUTO, steps: int = 20000,
                   wall_bc: int = _abi.BC_BOUNCE_BACK, marker_set: str = "fibonacci") -> Case:
    """Falling sphere with the immersed boundary: D3Q19 MRT on a uniform box, a sphere of the given
    diameter (cells) released at rest near the top and falling along -z under
    (1 - rho_f/rho_B) m g; multi-direct forcing with the 4-point Peskin kernel by default.  With
    nu = (tau - 1/2)/3 = 0.048 and g = 1e-4 the settling Reynolds number is about 10 (Schiller-
    Naumann drag).  The fluid itself carries no gravity (the buoyancy is in the body equation)."""
    nx, ny, nz = n
    pos, wid = zip(*[_uniform_axis(k) for k in n])
    R = diameter / 2.0
    pts = fibonacci_sphere(markers) if marker_set == "fibonacci" else icosphere(int(marker_set))
    spacing = (np.sqrt(4.0 * np.pi / len(pts)) * R if marker_set == "fibonacci"
               else 4.0 * R / np.sqrt(10.0 + 2.0 * np.sqrt(5.0)) / int(marker_set))
    centre = (nx / 2.0, ny / 2.0, nz - 3.0 * diameter)
    body = sphere_body(R, centre, pts, spacing, density_ratio, gravity=(0.0, 0.0, -gravity), method=method,
                       kernel=kernel, iterations=iterations)
    rates = np.full(27, 1.0)
    return Case(name=f"falling_sphere_{nx}x{ny}x{nz}", lattice=19, n=n, pos=list(pos), width=list(wid),
                periodic=(0, 0, 0), wall_bc=wall_bc, flux_axes=_abi.FLUX_NONE, stencil=5, s_rates=rates, tau=tau,
                storage=storage, path=path, steps=steps, body=body)


def cylinder_2d(n=(256, 256), diameter: float = None, markers: int = 400, density_ratio: float = 3.0,
                tau: float = 0.7, gravity: float = 5e-5, method: int = _abi.IB_STRESS,
                storage: int = _abi.F32, path: int = _abi.PATH_AUTO, steps: int = 2000) -> Case:
    """A falling cylinder in the x-z plane (D2Q9 MRT, uniform box, bounce-back walls): markers on a
    circle, body mass and moment of inertia per unit depth (rho_B pi R^2, m R^2 / 2), the stress
    method of the reference's 2D cylinder drivers by default."""
    nx, nz = n
    pos, wid = zip(*[_uniform_axis(k) for k in (nx, 1, nz)])
    d = diameter if diameter is not None else max(6.0, min(nx, nz) / 5.0)
    R = d / 2.0
    th = 2.0 * np.pi * (np.arange(markers) + 0.5) / markers
    unit = np.stack([np.cos(th), np.zeros_like(th), np.sin(th)], axis=1)
    m = density_ratio * np.pi * R * R
    body = Body(markers=unit * R, normals=unit.copy(), spacing=2.0 * np.pi * R / markers, mass=m,
                density=density_ratio, inertia=(0.5 * m * R * R,) * 3, pos=(nx / 2.0, 0.5, nz / 2.0),
                gravity=(0.0, 0.0, -gravity), method=method)
    return Case(name=f"cylinder_{nx}x{nz}", lattice=9, n=(nx, 1, nz), pos=list(pos), width=list(wid),
                periodic=(0, 0, 0), wall_bc=_abi.BC_BOUNCE_BACK, flux_axes=_abi.FLUX_NONE, stencil=5,
                s_rates=np.array([1.0, 1.4, 1.4, 1.0, 1.2, 1.0, 1.2, 1 / tau, 1 / tau]), tau=tau,
                storage=storage, path=path, steps=steps, body=body)


def _uniform_axis(n: int, dz: float = 1.0):
    return axis_from_widths(np.full(n, dz))


def lattice_taylor_green(nx: int = 256, nz: int = 128, ratio: float = 1.01, tau: float = 0.6,
                         amp: float = 0.01, seed: int = SEED_BASE + 1, steps: int = 1000,
                         stencil: int = 5, storage: int = _abi.F64, path: int = _abi.PATH_AUTO) -> Case:
    """Plain-method case: D2Q9 MRT, x geometrically stretched from the centre, uniform z,
    periodic x, half-way bounce-back walls in z, rho = 1 + U(-1e-3, 1e-3) and a Taylor-Green
    velocity of amplitude `amp` that vanishes normal to the walls."""
    px, wx = axis_from_widths(stretched_widths(nx, ratio))
    py, wy = _uniform_axis(1)
    pz, wz = _uniform_axis(nz)
    Lx, Lz = wx.sum(), wz.sum()
    case = Case(name=f"taylor_green_{nx}x{nz}", lattice=9, n=(nx, 1, nz), pos=[px, py, pz],
                width=[wx, wy, wz], periodic=(1, 0, 0), wall_bc=_abi.BC_BOUNCE_BACK,
                flux_axes=_abi.FLUX_X, stencil=stencil,
                s_rates=np.array([0, 1.4, 1.4, 0, 1.2, 0, 1.2, 1 / tau, 1 / tau]), tau=tau,
                storage=storage, path=path, steps=steps)
    X, _, Z = case.coords()
    kx, kz = 2 * np.pi / Lx, np.pi / Lz
    rho = 1.0 + uniform(seed, case.n_cells, -1e-3, 1e-3)
    case.init = {
        "p": rho - 1.0,
        "u": amp * np.sin(kx * X) * np.cos(kz * Z),
        "w": -amp * (kx / kz) * np.cos(kx * X) * np.sin(kz * Z),
    }
    return case


def lock_exchange(nx: int = 512, ny: int = 32, nz: int = 64, ratio: float = 1.02, tau: float = 0.55,
                  g_reduced: float = 1e-5, diffusivity: float = 1e-3, width: float = 3.0,
                  noise: float = 1e-4, seed: int = SEED_BASE + 2, steps: int = 5000,
                  storage: int = _abi.F64, path: int = _abi.PATH_AUTO, scalar_lattice: int = 7,
                  lattice: int = 19) -> Case:
    """Stratified lock exchange: D3Q19 MRT flow + D3Q7 salinity, Boussinesq buoyancy with
    reduced gravity g' = beta_s * g, z stretched toward mid-depth, slip walls at the x ends and
    in z, periodic span (y).  S = (1 - tanh((x - x0)/width))/2 across the lock gate x0 = Lx/2
    plus seeded noise of amplitude `noise` where 0.01 < S < 0.99."""
    px, wx = _uniform_axis(nx)
    py, wy = _uniform_axis(ny)
    pz, wz = axis_from_widths(stretched_widths(nz, ratio))
    cs2 = 0.25 if scalar_lattice == 7 else 1.0 / 3.0
    taus = diffusivity / cs2 + 0.5
    beta_s = 1e-3
    rates = np.ones(27)
    case = Case(name=f"lock_exchange_{nx}x{ny}x{nz}", lattice=lattice, n=(nx, ny, nz),
                pos=[px, py, pz], width=[wx, wy, wz], periodic=(0, 1, 0), wall_bc=_abi.BC_SLIP,
                flux_axes=_abi.FLUX_Z, stencil=5, s_rates=rates, tau=tau, taus=taus,
                rho_L=1.0, rho_H=1.0, beta_s=beta_s, gravity=g_reduced / beta_s, boussinesq=1,
                scalar=_abi.SCALAR_SALINITY, scalar_lattice=scalar_lattice, storage=storage,
                path=path, steps=steps)
    X, _, _ = case.coords()
    x0 = wx.sum() / 2.0
    S = 0.5 * (1.0 - np.tanh((X - x0) / width))
    band = (S > 0.01) & (S < 0.99)
    S = S + band * uniform(seed, case.n_cells, -noise, noise)
    case.init = {"S": S}
    return case


def solitary_wave(nx: int = 1024, nz: int = 256, ratio: float = 1.01, width: float = 4.0,
                  mobility: float = 0.02, rho_ratio: float = 1.03, depth_lower: float = 0.6,
                  amp: float = 0.3, x0: float = 200.0, noise: float = 1e-5, tau: float = 0.6,
                  gravity: float = 1e-5, sigma: float = 0.0, seed: int = SEED_BASE + 4,
                  steps: int = 10000, storage: int = _abi.F64, path: int = _abi.PATH_AUTO) -> Case:
    """Two-layer internal solitary wave: D2Q9 MRT flow + D2Q9 conservative Allen-Cahn phase
    field (collision.cu:190-223; step order of J_Zhou_solitary_wave4B1/Zhou_soliton4B1.cu:220-277).
    phi = 1 marks the denser lower layer (rho_H = rho_ratio * rho_L), z is stretched away from
    the undisturbed interface (finest there), the interface is a depression sech^2 hump of
    amplitude `amp` times the upper-layer depth centred at x0, plus seeded noise on phi."""
    px, wx = _uniform_axis(nx)
    py, wy = _uniform_axis(1)
    # finest cells at the interface height, geometric growth toward bottom and top
    i_int = int(round(depth_lower * nz))
    j = np.abs(np.arange(nz) - (i_int - 0.5))
    wz = ratio ** np.floor(j)
    pz, wz = axis_from_widths(wz)
    Lz = wz.sum()
    z_int = np.cumsum(wz)[i_int - 1]
    h_upper = Lz - z_int
    case = Case(name=f"solitary_wave_{nx}x{nz}", lattice=9, n=(nx, 1, nz), pos=[px, py, pz],
                width=[wx, wy, wz], periodic=(1, 0, 0), wall_bc=_abi.BC_SLIP, flux_axes=_abi.FLUX_Z,
                stencil=5, s_rates=np.array([0, 1.4, 1.4, 0, 1.2, 0, 1.2, 1 / tau, 1 / tau]), tau=tau,
                taus=0.5 + 3 * mobility, pf_tau=0.5 + 3 * mobility, rho_L=1.0, rho_H=rho_ratio,
                gravity=gravity, boussinesq=0, sigma=sigma, pf_thickness=width,
                scalar=_abi.SCALAR_PHASE_FIELD, storage=storage, path=path, steps=steps)
    X, _, Z = case.coords()
    Delta = 20.0 * width
    eta = -amp * h_upper / np.cosh((X - x0) / Delta) ** 2
    phi = 0.5 - 0.5 * np.tanh(2.0 * (Z - (z_int + eta)) / width)
    phi = phi + uniform(seed, case.n_cells, -noise, noise)
    case.init = {"phi": phi}
    return case


def phase_field_3d(n=(20, 6, 16), ratio: float = 1.02, width: float = 3.0, mobility: float = 0.02,
                   rho_ratio: float = 1.05, sigma: float = 1e-3, gravity: float = 1e-5, seed: int = SEED_BASE + 21,
                   storage: int = _abi.F64, path: int = _abi.PATH_AUTO) -> Case:
    """A two-layer interface in 3D: D3Q19 MRT flow + D3Q19 conservative Allen-Cahn phase field
    (the 3D form of collision.cu:190-223, whose drivers are 2D), z stretched toward mid-depth
    (the lock exchange's grid), periodic in x, slip walls, the interface tilted in x and y, surface
    tension on.  phi = 1 marks the denser lower layer."""
    c = lock_exchange(nx=n[0], ny=n[1], nz=n[2], ratio=ratio, storage=storage, path=path, scalar_lattice=0)
    c = c.replace(name=f"phase_field_3d_{n[0]}x{n[1]}x{n[2]}", scalar=_abi.SCALAR_PHASE_FIELD, periodic=(1, 0, 0),
                  taus=0.5 + 3 * mobility, pf_tau=0.5 + 3 * mobility, rho_L=1.0, rho_H=rho_ratio, gravity=gravity,
                  boussinesq=0, sigma=sigma, pf_thickness=width)
    X, Y, Z = c.coords()
    zi = 0.5 * Z.max() + 1.5 * np.sin(2 * np.pi * X / n[0]) + 0.5 * np.cos(2 * np.pi * Y / n[1])
    c.init = {"phi": 0.5 - 0.5 * np.tanh(2.0 * (Z - zi) / width), "p": uniform(seed, c.n_cells, -1e-4, 1e-4)}
    return c


@dataclasses.dataclass
class Items:
    """The run constants a driver reads from its input/initial.csv and derives from them
    (input_items and Items::setc, function/basic.cu:57-98)."""
    nx: int
    ny: int
    nz: int
    dx: float
    dt: float
    ratiox: float
    ratioy: float
    gz: float
    nu: float              # kinematic viscosity, m^2/s (the file holds it in 1e-6 m^2/s)
    save_interval: int
    total_count: int       # steps: 1 + save_interval * (the file's count of saves)
    num_velocity: int      # 9 when the grid is one cell deep in y, else 27
    c: float               # dx / dt
    tau: float             # 3e-6 / (c^2 dt) + 1/2: setc fixes nu = 1e-6 here, not the file's nu
    taus: float            # 1.4 * 3 * 10^-alpha / (c^2 dt) + 1/2

    def case_fields(self) -> dict:
        """The ``Case`` fields these constants set (the driver-specific geometry, widths and
        fields are not in the file)."""
        return {"lattice": self.num_velocity, "n": (self.nx, self.ny, self.nz), "dz": self.dx,
                "dt": self.dt, "tau": self.tau, "taus": self.taus, "wall_nu": self.nu,
                "steps": self.total_count}


def read_initial_csv(path: str, alpha: int = 9) -> Items:
    """Parses a driver's input/initial.csv (its first line: nx, ny, nz, dx, dt, ratiox, ratioy,
    gz, nu, save_interval, total_count) and applies input_items' and setc(alpha)'s derivations
    (basic.cu:57-77; the drivers call setc(9), e.g. LockExchange.cu:16)."""
    with open(path) as fh:
        first = fh.readline()
    v = [x.strip() for x in first.split(",")]
    nx, ny, nz = int(v[0]), int(v[1]), int(v[2])
    dx, dt, ratiox, ratioy, gz, nu = (float(x) for x in v[3:9])
    save_interval, count = int(v[9]), int(v[10])
    total = 1 + save_interval * count
    nu *= 10.0 ** -6
    nx = int(nx / ratiox)
    ny = int(ny / ratioy)
    if ny < 1:
        ny, ratioy = 1, 1.0
    if ny == 1:
        q, ratioy = 9, 1.0
    else:
        q = 27
    c = dx / dt
    tau = 1.0 * 3.0 * 10.0 ** -6 / c / c / dt + 0.5
    taus = 1.4 * 3.0 * 10.0 ** -alpha / c / c / dt + 0.5
    return Items(nx=nx, ny=ny, nz=nz, dx=dx, dt=dt, ratiox=ratiox, ratioy=ratioy, gz=gz, nu=nu,
                 save_interval=save_interval, total_count=total, num_velocity=q, c=c, tau=tau, taus=taus)


def read_rates_csv(path: str, q: int) -> np.ndarray:
    """The relaxation rates of a driver's input/S_D3Q{q}.csv as ``Case.s_rates`` (set_M,
    basic.cu:171-189): the first q values of the first line, with the stress rows zeroed as set_M
    zeroes them (rows 7, 8 of D2Q9; 5-9 of D3Q27), since they relax at 1/tau of each cell."""
    with open(path) as fh:
        vals = [float(x) for x in fh.readline().split(",")[:q] if x.strip()]
    if len(vals) != q:
        raise ValueError(f"{path}: {len(vals)} rates, expected {q}")
    r = np.array(vals)
    if q == 9:
        r[7] = r[8] = 0.0
    elif q == 27:
        r[5:10] = 0.0
    return r


CASES = {
    "taylor_green": lattice_taylor_green,
    "lock_exchange": lock_exchange,
    "solitary_wave": solitary_wave,
}
