"""Seeded synthetic initial conditions (host numpy), the inputs of the PM benchmarks.

- ``plummer_ic``: an isolated Plummer sphere in equilibrium — positions from the inverse of
  the cumulative mass M(<r)/M = r^3/(r^2+a^2)^{3/2}, speeds q*v_esc with q drawn by
  rejection from q^2(1-q^2)^{7/2} under a 0.1 envelope, isotropic directions, centre of
  mass and mean velocity removed (the recipe of reference ic.py:296-380).
- ``two_galaxy_ic``: the collision recipe of the benchmark configurations — per galaxy an
  exponential disk (scale length r_d, sech^2 vertical profile of height z0, 80% of the
  stellar mass), a Plummer bulge (scale a_b, 20%) and a central black-hole particle (1% of
  the stellar mass); circular velocities from the enclosed mass plus Gaussian dispersion;
  disks inclined about the collision (x) axis; galaxies placed symmetrically about the box
  centre along x and closing at a fraction of the mutual escape speed.

Distances are in the box's length unit with G = 1 unless given.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

COMPONENT_DISK, COMPONENT_BULGE, COMPONENT_BH = 0, 1, 2


@dataclass
class ICResult:
    pos: np.ndarray        # (N, 3) float32
    vel: np.ndarray        # (N, 3) float32
    mass: np.ndarray       # (N,) float32
    gid: np.ndarray        # (N,) int32 galaxy id
    component: np.ndarray  # (N,) int8
    preset: str

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])


def _unit_vectors(rng, n: int) -> np.ndarray:
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v


def _plummer_speed_ratio(rng, n: int) -> np.ndarray:
    out = np.empty(n)
    filled = 0
    while filled < n:
        m = n - filled
        x = rng.random(m)
        y = rng.random(m)
        ok = 0.1 * y < x ** 2 * (1.0 - x ** 2) ** 3.5
        k = int(ok.sum())
        out[filled:filled + k] = x[ok]
        filled += k
    return out


def _plummer_sample(rng, n: int, mass: float, a: float, grav: float, rmax_factor: float = 20.0):
    u = np.clip(rng.random(n), 1e-12, 1.0)
    r = np.minimum(a / np.sqrt(u ** (-2.0 / 3.0) - 1.0), rmax_factor * a)
    pos = r[:, None] * _unit_vectors(rng, n)
    q = _plummer_speed_ratio(rng, n)
    v_esc = np.sqrt(2.0 * grav * mass) * (r ** 2 + a ** 2) ** -0.25
    vel = (q * v_esc)[:, None] * _unit_vectors(rng, n)
    return pos, vel


def plummer_ic(n: int, *, total_mass: float = 1.0, scale_a: float = 1.0, grav: float = 1.0,
               centre=(0.0, 0.0, 0.0), seed: int = 0) -> ICResult:
    """Equal-mass Plummer sphere at `centre`, at rest."""
    rng = np.random.default_rng(seed)
    pos, vel = _plummer_sample(rng, n, total_mass, scale_a, grav)
    pos -= pos.mean(axis=0)
    vel -= vel.mean(axis=0)
    pos += np.asarray(centre, dtype=np.float64)
    return ICResult(pos.astype(np.float32), vel.astype(np.float32),
                    np.full(n, total_mass / n, np.float32), np.zeros(n, np.int32),
                    np.full(n, COMPONENT_BULGE, np.int8), "plummer")


def _exp_disk_radius(u: np.ndarray, r_d: float) -> np.ndarray:
    """Inverse of M(<R)/M_d = 1 - (1 + R/r_d) exp(-R/r_d) via the W_{-1} branch."""
    from scipy.special import lambertw

    w = lambertw(-(1.0 - u) / np.e, k=-1).real
    return r_d * (-1.0 - w)


def _rotate_x(v: np.ndarray, deg: float) -> np.ndarray:
    t = np.deg2rad(deg)
    c, s = np.cos(t), np.sin(t)
    out = v.copy()
    out[:, 1] = c * v[:, 1] - s * v[:, 2]
    out[:, 2] = s * v[:, 1] + c * v[:, 2]
    return out


def galaxy(rng, n: int, *, mass: float = 1.0, disk_frac: float = 0.8, r_d: float = 1.0,
           z0: float = 0.1, bulge_a: float = 0.2, bh_frac: float = 0.01,
           dispersion: float = 0.1, grav: float = 1.0, r_max: float = 9.0,
           softening: float = 0.05):
    """One galaxy centred at the origin in its own disk plane: (pos, vel, mass, component).

    n stellar particles (disk + bulge, equal masses) plus one black-hole particle.
    """
    n_disk = int(round(disk_frac * n))
    n_bulge = n - n_disk
    m_disk, m_bulge = mass * disk_frac, mass * (1.0 - disk_frac)
    m_bh = bh_frac * mass
    # disk: radius from the exponential profile truncated at r_max, sech^2 heights
    u_max = 1.0 - (1.0 + r_max / r_d) * np.exp(-r_max / r_d)
    R = _exp_disk_radius(rng.random(n_disk) * u_max, r_d)
    phi = rng.uniform(0.0, 2.0 * np.pi, n_disk)
    z = z0 * np.arctanh(np.clip(2.0 * rng.random(n_disk) - 1.0, -1 + 1e-12, 1 - 1e-12))
    d_pos = np.stack([R * np.cos(phi), R * np.sin(phi), z], axis=1)
    x = R / r_d
    m_enc = (m_disk * (1.0 - (1.0 + x) * np.exp(-x)) / u_max
             + m_bulge * R ** 3 / (R ** 2 + bulge_a ** 2) ** 1.5 + m_bh)
    r_eff = np.sqrt(R ** 2 + softening ** 2)
    v_c = np.sqrt(grav * m_enc * R ** 2 / r_eff ** 3)
    tang = np.stack([-np.sin(phi), np.cos(phi), np.zeros_like(phi)], axis=1)
    d_vel = v_c[:, None] * tang + dispersion * v_c[:, None] * rng.normal(size=(n_disk, 3))
    # bulge: Plummer sphere with its isotropic dispersion sigma^2 = G M_b / (6 sqrt(r^2+a^2))
    b_pos, _ = _plummer_sample(rng, n_bulge, m_bulge, bulge_a, grav, rmax_factor=r_max / bulge_a)
    rb = np.linalg.norm(b_pos, axis=1)
    sigma = np.sqrt(grav * (m_bulge / (6.0 * np.sqrt(rb ** 2 + bulge_a ** 2))
                            + m_bh / (3.0 * np.sqrt(rb ** 2 + softening ** 2))))
    b_vel = sigma[:, None] * rng.normal(size=(n_bulge, 3))
    pos = np.concatenate([d_pos, b_pos, np.zeros((1, 3))])
    vel = np.concatenate([d_vel, b_vel, np.zeros((1, 3))])
    m = np.concatenate([np.full(n_disk, m_disk / n_disk), np.full(n_bulge, m_bulge / n_bulge), [m_bh]])
    comp = np.concatenate([np.full(n_disk, COMPONENT_DISK), np.full(n_bulge, COMPONENT_BULGE),
                           [COMPONENT_BH]]).astype(np.int8)
    return pos, vel, m, comp


def two_galaxy_ic(n_per_galaxy, *, box_centre, separation: float = 20.0,
                  approach: float = 0.5, inclinations=(30.0, -45.0), masses=(1.0, 1.0),
                  grav: float = 1.0, seed: int = 0, **galaxy_kw) -> ICResult:
    """Two galaxies on a head-on approach along x (see the module docstring).

    n_per_galaxy: int or (n1, n2) stellar particles per galaxy (each also gets a BH).
    masses: stellar masses; the closing speed is `approach` times sqrt(2 G M_tot / sep)
    with M_tot including the black holes.
    """
    rng = np.random.default_rng(seed)
    counts = (n_per_galaxy, n_per_galaxy) if np.isscalar(n_per_galaxy) else tuple(n_per_galaxy)
    bh_frac = galaxy_kw.get("bh_frac", 0.01)
    m_tot = sum(masses) * (1.0 + bh_frac)
    v_rel = approach * np.sqrt(2.0 * grav * m_tot / separation)
    centre = np.asarray(box_centre, dtype=np.float64)
    parts = []
    for g, (n, m, inc) in enumerate(zip(counts, masses, inclinations)):
        pos, vel, mm, comp = galaxy(rng, int(n), mass=m, grav=grav, **galaxy_kw)
        pos, vel = _rotate_x(pos, inc), _rotate_x(vel, inc)
        # galaxy 0 on the left moving +x, galaxy 1 on the right moving -x; the bulk speeds
        # split v_rel so that the total momentum is zero
        m_this, m_other = m, masses[1 - g]
        side = -1.0 if g == 0 else 1.0
        offset = separation * m_other / (m_this + m_other)
        speed = v_rel * m_other / (m_this + m_other)
        pos = pos + centre + np.array([side * offset, 0.0, 0.0])
        vel = vel + np.array([-side * speed, 0.0, 0.0])
        parts.append((pos, vel, mm, np.full(len(mm), g, np.int32), comp))
    pos, vel, mass, gid, comp = (np.concatenate([p[i] for p in parts]) for i in range(5))
    return ICResult(pos.astype(np.float32), vel.astype(np.float32), mass.astype(np.float32),
                    gid, comp, "two_galaxy")


# ---------------------------------------------------------------------------------------
# the same collision recipe drawn with torch, on the device (100M-particle inputs in seconds)
# ---------------------------------------------------------------------------------------

@dataclass
class TensorIC:
    """Initial conditions as torch tensors (on the device they were drawn on)."""

    pos: "torch.Tensor"        # (N, 3) float32
    vel: "torch.Tensor"        # (N, 3) float32
    mass: "torch.Tensor"       # (N,) float32
    gid: "torch.Tensor"        # (N,) int32
    component: "torch.Tensor"  # (N,) int8
    preset: str

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])

    def to_host(self) -> ICResult:
        return ICResult(self.pos.cpu().numpy(), self.vel.cpu().numpy(), self.mass.cpu().numpy(),
                        self.gid.cpu().numpy(), self.component.cpu().numpy(), self.preset)


def _exp_disk_radius_t(u, r_d: float, x_max: float):
    """R with 1 - (1 + R/r_d) exp(-R/r_d) = u for u in [0, u(x_max)]: bisection in float64 on
    [0, x_max] (the cumulative mass is increasing; 60 halvings reach float64 resolution)."""
    import torch

    lo = torch.zeros_like(u)
    hi = torch.full_like(u, x_max)
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        below = 1.0 - (1.0 + mid) * torch.exp(-mid) < u
        lo = torch.where(below, mid, lo)
        hi = torch.where(below, hi, mid)
    return r_d * 0.5 * (lo + hi)


def galaxy_tensors(gen, n: int, *, mass: float = 1.0, disk_frac: float = 0.8,
                   r_d: float = 1.0, z0: float = 0.1, bulge_a: float = 0.2,
                   bh_frac: float = 0.01, dispersion: float = 0.1, grav: float = 1.0,
                   r_max: float = 9.0, softening: float = 0.05, device=None):
    """``galaxy`` drawn with the torch generator ``gen`` (float64 arithmetic): the same
    distributions, not the same random stream. Returns (pos, vel, mass, component)."""
    import torch

    f64 = dict(dtype=torch.float64, device=device)
    n_disk = int(round(disk_frac * n))
    n_bulge = n - n_disk
    m_disk, m_bulge, m_bh = mass * disk_frac, mass * (1.0 - disk_frac), bh_frac * mass
    u_max = 1.0 - (1.0 + r_max / r_d) * np.exp(-r_max / r_d)
    R = _exp_disk_radius_t(torch.rand(n_disk, generator=gen, **f64) * u_max, r_d, r_max / r_d)
    phi = torch.rand(n_disk, generator=gen, **f64) * (2.0 * np.pi)
    z = z0 * torch.atanh(torch.clamp(2.0 * torch.rand(n_disk, generator=gen, **f64) - 1.0,
                                     -1 + 1e-12, 1 - 1e-12))
    cphi, sphi = torch.cos(phi), torch.sin(phi)
    d_pos = torch.stack([R * cphi, R * sphi, z], 1)
    x = R / r_d
    m_enc = (m_disk * (1.0 - (1.0 + x) * torch.exp(-x)) / u_max
             + m_bulge * R ** 3 / (R ** 2 + bulge_a ** 2) ** 1.5 + m_bh)
    v_c = torch.sqrt(grav * m_enc * R ** 2 / (R ** 2 + softening ** 2) ** 1.5)
    tang = torch.stack([-sphi, cphi, torch.zeros_like(phi)], 1)
    d_vel = v_c[:, None] * (tang + dispersion * torch.randn(n_disk, 3, generator=gen, **f64))
    del R, phi, z, cphi, sphi, x, m_enc, tang
    # bulge: Plummer positions from the inverse cumulative mass, Gaussian velocities with
    # the isotropic dispersion of the bulge plus the black hole
    u = torch.clamp(torch.rand(n_bulge, generator=gen, **f64), 1e-12, 1.0)
    rb = torch.clamp(bulge_a / torch.sqrt(u ** (-2.0 / 3.0) - 1.0), max=r_max)
    dirs = torch.randn(n_bulge, 3, generator=gen, **f64)
    dirs /= torch.linalg.vector_norm(dirs, dim=1, keepdim=True)
    b_pos = rb[:, None] * dirs
    sigma = torch.sqrt(grav * (m_bulge / (6.0 * torch.sqrt(rb ** 2 + bulge_a ** 2))
                               + m_bh / (3.0 * torch.sqrt(rb ** 2 + softening ** 2))))
    b_vel = sigma[:, None] * torch.randn(n_bulge, 3, generator=gen, **f64)
    del u, rb, dirs, sigma
    zero = torch.zeros(1, 3, **f64)
    pos = torch.cat([d_pos, b_pos, zero])
    vel = torch.cat([d_vel, b_vel, zero])
    m = torch.cat([torch.full((n_disk,), m_disk / n_disk, **f64),
                   torch.full((n_bulge,), m_bulge / n_bulge, **f64),
                   torch.full((1,), m_bh, **f64)])
    comp = torch.cat([torch.full((n_disk,), COMPONENT_DISK, dtype=torch.int8, device=device),
                      torch.full((n_bulge,), COMPONENT_BULGE, dtype=torch.int8, device=device),
                      torch.full((1,), COMPONENT_BH, dtype=torch.int8, device=device)])
    return pos, vel, m, comp


def _rotate_x_t(v, deg: float):
    t = np.deg2rad(deg)
    c, s = float(np.cos(t)), float(np.sin(t))
    out = v.clone()
    out[:, 1] = c * v[:, 1] - s * v[:, 2]
    out[:, 2] = s * v[:, 1] + c * v[:, 2]
    return out


def two_galaxy_ic_tensors(n_per_galaxy, *, box_centre, separation: float = 20.0,
                          approach: float = 0.5, inclinations=(30.0, -45.0), masses=(1.0, 1.0),
                          grav: float = 1.0, seed: int = 0, device=None,
                          sort_cells: tuple[int, float] | None = None, **galaxy_kw) -> TensorIC:
    """``two_galaxy_ic`` drawn with torch on ``device`` (same recipe, placement and bulk
    velocities; its own random stream). ``sort_cells=(grid_size, dx)`` also orders the
    particles by cell as ``sort_by_cell`` does."""
    import torch

    gen = torch.Generator(device=device or "cpu")
    gen.manual_seed(int(seed))
    counts = (n_per_galaxy, n_per_galaxy) if np.isscalar(n_per_galaxy) else tuple(n_per_galaxy)
    bh_frac = galaxy_kw.get("bh_frac", 0.01)
    v_rel = approach * np.sqrt(2.0 * grav * sum(masses) * (1.0 + bh_frac) / separation)
    centre = torch.tensor(box_centre, dtype=torch.float64, device=device)
    out = {k: [] for k in ("pos", "vel", "mass", "gid", "comp")}
    for g, (n, m, inc) in enumerate(zip(counts, masses, inclinations)):
        pos, vel, mm, comp = galaxy_tensors(gen, int(n), mass=m, grav=grav, device=device,
                                            **galaxy_kw)
        pos, vel = _rotate_x_t(pos, inc), _rotate_x_t(vel, inc)
        m_other = masses[1 - g]
        side = -1.0 if g == 0 else 1.0
        pos += centre
        pos[:, 0] += side * separation * m_other / (m + m_other)
        vel[:, 0] += -side * v_rel * m_other / (m + m_other)
        out["pos"].append(pos.float())
        out["vel"].append(vel.float())
        out["mass"].append(mm.float())
        out["gid"].append(torch.full((mm.shape[0],), g, dtype=torch.int32, device=device))
        out["comp"].append(comp)
        del pos, vel, mm
    ic = TensorIC(*(torch.cat(out[k]) for k in ("pos", "vel", "mass", "gid", "comp")),
                  preset="two_galaxy")
    if sort_cells is not None:
        grid_size, dx = sort_cells
        g = torch.clamp(torch.floor(ic.pos.double() / dx).long(), -1, grid_size)
        mm = grid_size + 2
        key = ((g[:, 0] + 1) * mm + (g[:, 1] + 1)) * mm + (g[:, 2] + 1)
        del g
        order = torch.sort(key, stable=True).indices
        del key
        ic = TensorIC(ic.pos[order], ic.vel[order], ic.mass[order], ic.gid[order],
                      ic.component[order], ic.preset)
    return ic


# ---------------------------------------------------------------------------------------
# disks launched on the rotation curve the grid produces (reference ic.py:408-525)
# ---------------------------------------------------------------------------------------

def equilibrate_disk_velocities(icr, *, dx: float, solver, rho, phi, ax, ay, az,
                                n_radii: int = 64, n_azimuth: int = 64, parts=None,
                                dispersion: float = 0.0, seed: int = 0) -> dict:
    """Give every disk particle the circular speed of the force the PM chain will apply.

    For each galaxy its own particles alone are deposited on the production grid, solved with
    the production ``solver`` (cold) and differentiated; rings of probes in the disk plane
    through the galaxy's black hole sample the azimuthally averaged inward radial
    acceleration <-a_R>(R), so v_c(R) = sqrt(R <-a_R>) is the circular speed the grid
    supports (the clamp to >= 0 is applied to the mean, not per sample). Each disk particle
    gets v = R*omega(R) along n x R_hat plus the galaxy's bulk velocity, omega = v_c/R
    interpolated so that particles inside the first ring rotate as a solid body; with
    ``dispersion`` > 0 a Gaussian of that fraction of v_c is added per component (the
    benchmark recipe's 10%). Bulge and black-hole velocities are untouched. The disk plane
    is the one normal to the disk's angular momentum about the black hole (the reference's
    disks lie in the xy plane; the benchmark's are inclined).

    ``icr``: ICResult (numpy) or TensorIC (torch) with ``component``; its velocities are
    replaced; with ``parts`` (a ParticleState holding the same particles, in load order or
    reordered with ``pid``) the device velocities are written too. ``rho``..``az`` are
    scratch grid fields (the production ones work: a run re-primes them). Returns
    {gid: (r_table, vc_table)}.
    """
    import torch

    from galaxy_collision_amd.data import ParticleState, as_tensor
    from galaxy_collision_amd.deposit import (deposit_density, gather_acceleration,
                                              potential_to_accel)

    if getattr(icr, "component", None) is None:
        raise ValueError("equilibrate_disk_velocities needs the component tags of a "
                         "two-galaxy IC")
    dev = as_tensor(rho).device
    as_t = lambda a: (a if isinstance(a, torch.Tensor) else torch.from_numpy(np.asarray(a)))  # noqa: E731
    pos = as_t(icr.pos).to(dev, torch.float64)
    vel = as_t(icr.vel).to(dev, torch.float64)
    mass = as_t(icr.mass).to(dev, torch.float32)
    gid = as_t(icr.gid).to(dev)
    comp = as_t(icr.component).to(dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(seed))
    curves: dict = {}
    for g in torch.unique(gid).tolist():
        sel = torch.nonzero(gid == g).squeeze(1)
        bh = sel[comp[sel] == COMPONENT_BH]
        if bh.numel():
            centre, bulk = pos[bh[0]].clone(), vel[bh[0]].clone()
        else:  # no black hole: the stellar centroid, at rest
            centre, bulk = pos[sel].mean(0), torch.zeros(3, dtype=torch.float64, device=dev)
        disk = sel[comp[sel] == COMPONENT_DISK]
        if disk.numel() == 0:
            continue
        rel = pos[disk] - centre
        L = (mass[disk].double()[:, None] * torch.linalg.cross(rel, vel[disk] - bulk)).sum(0)
        nrm = L / torch.linalg.vector_norm(L)
        rvec = rel - (rel @ nrm)[:, None] * nrm
        R = torch.linalg.vector_norm(rvec, dim=1)
        del rel
        # this galaxy's self-gravity on the production grid
        sub = ParticleState(int(sel.numel()), device=dev)
        sub.load_tensors(pos[sel].float(), None, mass[sel])
        deposit_density(sub, rho, dx)
        solver.solve(rho, phi, warm_start=False)
        potential_to_accel(phi, ax, ay, az, dx)
        del sub
        # probe rings in the disk plane
        helper = torch.tensor([1.0, 0.0, 0.0] if abs(float(nrm[0])) < 0.9 else [0.0, 1.0, 0.0],
                              dtype=torch.float64, device=dev)
        e1 = torch.linalg.cross(nrm, helper)
        e1 /= torch.linalg.vector_norm(e1)
        e2 = torch.linalg.cross(nrm, e1)
        r_tab = torch.linspace(float(dx), float(R.max()), n_radii, dtype=torch.float64,
                               device=dev)
        theta = torch.arange(n_azimuth, dtype=torch.float64, device=dev) * (2 * np.pi / n_azimuth)
        dirs = torch.cos(theta)[:, None] * e1 + torch.sin(theta)[:, None] * e2  # (az, 3)
        ppos = centre + r_tab[:, None, None] * dirs[None]                        # (r, az, 3)
        probes = ParticleState(n_radii * n_azimuth, device=dev)
        probes.load_tensors(ppos.reshape(-1, 3).float())
        acc = [torch.empty(probes.n, device=dev) for _ in range(3)]
        gather_acceleration(probes, ax, ay, az, *acc, dx)
        a = torch.stack(acc, 1).double().reshape(n_radii, n_azimuth, 3)
        a_rad = (a * dirs[None]).sum(-1)
        inward = torch.clamp(-a_rad.mean(1), min=0.0)
        vc_tab = torch.sqrt(r_tab * inward)
        curves[int(g)] = (r_tab.cpu().numpy(), vc_tab.cpu().numpy())
        # omega = v_c/R interpolated linearly in R, held at the end values outside the table
        omega_tab = vc_tab / r_tab
        k = torch.clamp(torch.searchsorted(r_tab, R), 1, n_radii - 1)
        r0, r1 = r_tab[k - 1], r_tab[k]
        t = torch.clamp((R - r0) / (r1 - r0), 0.0, 1.0)
        omega = omega_tab[k - 1] + t * (omega_tab[k] - omega_tab[k - 1])
        vc = R * omega
        that = torch.linalg.cross(nrm.expand_as(rvec), rvec / torch.clamp(R, min=1e-6)[:, None])
        v_new = vc[:, None] * that
        if dispersion > 0.0:
            v_new += dispersion * vc[:, None] * torch.randn(v_new.shape, generator=gen,
                                                            dtype=torch.float64, device=dev)
        vel[disk] = v_new + bulk
        del rvec, R, vc, that, v_new
    vel32 = vel.float()
    if isinstance(icr.vel, torch.Tensor):
        icr.vel = vel32.to(icr.vel.device)
    else:
        icr.vel = vel32.cpu().numpy()
    if parts is not None:
        pid = parts.pid.tensor.long()
        for a_, f in enumerate((parts.vel_x, parts.vel_y, parts.vel_z)):
            f.tensor.copy_(vel32[pid, a_])
    return curves


def sort_by_cell(ic: ICResult, grid_size: int, dx: float) -> ICResult:
    """Reorders particles by their CIC cell (C order): the memory layout the fused step
    prefers (neighbouring threads touch neighbouring nodes). Particle identity is kept in
    ``gid``/``component``; the physics does not depend on the order."""
    g = np.floor(ic.pos.astype(np.float64) / dx).astype(np.int64)
    g = np.clip(g, -1, grid_size)
    m = grid_size + 2
    key = ((g[:, 0] + 1) * m + (g[:, 1] + 1)) * m + (g[:, 2] + 1)
    order = np.argsort(key, kind="stable")
    return ICResult(ic.pos[order], ic.vel[order], ic.mass[order], ic.gid[order],
                    ic.component[order], ic.preset)


def box_fit_fraction(ic: ICResult, grid_size: int, dx: float) -> float:
    """Mass fraction outside the node span [0, (grid_size-1)*dx] (reference sim.py:173-191)."""
    hi = (grid_size - 1) * dx
    outside = np.any((ic.pos < 0.0) | (ic.pos > hi), axis=1)
    return float(ic.mass[outside].astype(np.float64).sum() / ic.mass.astype(np.float64).sum())


__all__ = ["ICResult", "TensorIC", "plummer_ic", "two_galaxy_ic", "two_galaxy_ic_tensors", "galaxy",
           "galaxy_tensors", "equilibrate_disk_velocities", "sort_by_cell", "box_fit_fraction"]
