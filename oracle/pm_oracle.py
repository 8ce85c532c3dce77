"""CPU oracle for the PM force chain — TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg import this
module, and only as the checker; the product library never calls it.

It restates the reference (pouyalj/Galaxy-Collision-with-CUDA-FFT, Taichi) in numpy float32
with the reference's per-element operation order, so single kernels can be compared bit for
bit, plus float64 ground truths that pin those restatements to mathematics:

- ``cell_keys`` / ``stable_perm``: integer cell keys and np.argsort(kind="stable");
- ``deposit_cic`` (C, ``pm_oracle.c``): the deterministic deposit order;
  ``deposit_cic_f64``: float64 np.add.at scatter of the same weights (ground truth);
- ``grad`` (deposit.py:132-174), ``gather`` (deposit.py:85-129), ``kick``/``drift``
  (integrator.py:27-60), ``direct_accel`` (integrator.py:63-97);
- multigrid pieces (solver/multigrid.py): ``rb_pass`` (46-59), ``residual`` (62-79),
  ``restrict`` (82-97), ``prolong_add`` (100-130), ``moments`` (391-418),
  ``face_params`` (378-389), ``set_faces`` (261-295), and the full PLAIN ``mg_solve``
  (455-487);
- ``dirichlet_solve_f64``: exact float64 DST-I solve of the same discrete Dirichlet problem;
- ``fft_solve_f64``: the zero-padded continuum-kernel FFT solve (solver/fft_oracle.py:82-109);
- diagnostics (diagnostics_device.py:41-87, 152-184; diagnostics.py:69-173).
"""

from __future__ import annotations

import ctypes
import functools
import math
from pathlib import Path

import numpy as np

F32 = np.float32
_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "libpmoracle.so"
_lib = None


def lib():
    """The compiled C half of the oracle (built by ``make -C oracle``)."""
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise RuntimeError(f"{_LIB_PATH} is missing: run `make -C oracle`")
        L = ctypes.CDLL(str(_LIB_PATH))
        fp = ctypes.POINTER(ctypes.c_float)
        i64 = ctypes.c_int64
        L.oracle_deposit_cic.argtypes = [fp, fp, fp, fp, i64, ctypes.c_int, ctypes.c_float, fp,
                                         ctypes.POINTER(ctypes.c_uint32),
                                         ctypes.POINTER(ctypes.c_int32)]
        L.oracle_deposit_cic.restype = ctypes.c_int
        L.oracle_deposit_fixed.argtypes = [fp, fp, fp, fp, i64, ctypes.c_int, ctypes.c_float, fp,
                                           ctypes.POINTER(ctypes.c_uint32)]
        L.oracle_deposit_fixed.restype = ctypes.c_int
        L.cpu_deposit_atomic.argtypes = [fp, fp, fp, fp, i64, ctypes.c_int, ctypes.c_float, fp]
        L.cpu_grad.argtypes = [fp, fp, fp, fp, ctypes.c_int, ctypes.c_float]
        L.cpu_gather.argtypes = [fp, fp, fp, fp, fp, fp, fp, fp, fp, i64, ctypes.c_int,
                                 ctypes.c_float]
        L.cpu_kick.argtypes = [fp, fp, fp, fp, fp, fp, i64, ctypes.c_float]
        L.cpu_drift.argtypes = [fp, fp, fp, fp, fp, fp, i64, ctypes.c_float]
        L.cpu_mg_create.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_double]
        L.cpu_mg_create.restype = ctypes.c_void_p
        L.cpu_mg_destroy.argtypes = [ctypes.c_void_p]
        L.cpu_mg_solve.argtypes = [ctypes.c_void_p, fp, fp, ctypes.c_int]
        L.cpu_mg_solve.restype = ctypes.c_int
        L.cpu_num_threads.restype = ctypes.c_int
        L.cpu_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _fp(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


# ---------------------------------------------------------------------------------------
# CIC geometry and deposit
# ---------------------------------------------------------------------------------------

def _cic_axis(x: np.ndarray, dx: float, n: int):
    """(in_range, base index, offset) per reference deposit.py:62-70 (float32)."""
    g = (x.astype(F32) / F32(dx)).astype(F32)
    ok = (g >= F32(-1.0)) & (g < F32(n))
    gi = np.where(ok, g, F32(0.0))
    i = np.floor(gi).astype(np.int32)
    f = (gi - i.astype(F32)).astype(F32)
    return ok, i, f


def cell_keys(pos: np.ndarray, n: int, dx: float) -> np.ndarray:
    """Extended cell key per particle, (n+1)^3 for particles with no in-box corner."""
    m = n + 1
    okx, ix, _ = _cic_axis(pos[:, 0], dx, n)
    oky, iy, _ = _cic_axis(pos[:, 1], dx, n)
    okz, iz, _ = _cic_axis(pos[:, 2], dx, n)
    key = ((ix.astype(np.int64) + 1) * m + (iy + 1)) * m + (iz + 1)
    ok = okx & oky & okz
    return np.where(ok, key, m ** 3).astype(np.uint32)


def stable_perm(keys: np.ndarray) -> np.ndarray:
    return np.argsort(keys, kind="stable").astype(np.int32)


def deposit_cic(pos: np.ndarray, mass: np.ndarray, n: int, dx: float):
    """The deterministic-order deposit (C). Returns (rho, keys, perm)."""
    pos = np.ascontiguousarray(pos, dtype=F32)
    mass = np.ascontiguousarray(mass, dtype=F32)
    x, y, z = (np.ascontiguousarray(pos[:, a]) for a in range(3))
    npart = mass.shape[0]
    rho = np.zeros((n, n, n), dtype=F32)
    keys = np.zeros(max(npart, 1), dtype=np.uint32)
    perm = np.zeros(max(npart, 1), dtype=np.int32)
    rc = lib().oracle_deposit_cic(
        _fp(x), _fp(y), _fp(z), _fp(mass), npart, n, dx, _fp(rho),
        keys.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
        perm.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    if rc != 0:
        raise MemoryError("oracle_deposit_cic allocation failed")
    return rho, keys[:npart], perm[:npart]


def deposit_cic_fixed(pos: np.ndarray, mass: np.ndarray, n: int, dx: float):
    """The fixed-point deposit (C). Returns (rho, keys)."""
    pos = np.ascontiguousarray(pos, dtype=F32)
    mass = np.ascontiguousarray(mass, dtype=F32)
    x, y, z = (np.ascontiguousarray(pos[:, a]) for a in range(3))
    npart = mass.shape[0]
    rho = np.zeros((n, n, n), dtype=F32)
    keys = np.zeros(max(npart, 1), dtype=np.uint32)
    rc = lib().oracle_deposit_fixed(
        _fp(x), _fp(y), _fp(z), _fp(mass), npart, n, dx, _fp(rho),
        keys.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    if rc != 0:
        raise MemoryError("oracle_deposit_fixed allocation failed")
    return rho, keys[:npart]


def deposit_cic_f64(pos: np.ndarray, mass: np.ndarray, n: int, dx: float) -> np.ndarray:
    """Float64 scatter of the same float32 CIC weights (order-free ground truth)."""
    rho = np.zeros(n ** 3, dtype=np.float64)
    inv_vol = F32(1.0) / (F32(dx) * F32(dx) * F32(dx))
    axes = [_cic_axis(pos[:, a], dx, n) for a in range(3)]
    ok = axes[0][0] & axes[1][0] & axes[2][0]
    m = (mass.astype(F32) * inv_vol).astype(np.float64)
    for c in range(8):
        bits = (c >> 2, (c >> 1) & 1, c & 1)
        w = np.ones_like(m)
        idx = np.zeros(m.shape, dtype=np.int64)
        inb = ok.copy()
        for a, b in enumerate(bits):
            _, i, f = axes[a]
            j = i.astype(np.int64) + b
            inb &= (j >= 0) & (j < n)
            w = w * (f.astype(np.float64) if b else (1.0 - f.astype(np.float64)))
            idx = idx * n + np.clip(j, 0, n - 1)
        np.add.at(rho, idx[inb], (m * w)[inb])
    return rho.reshape(n, n, n)


# ---------------------------------------------------------------------------------------
# gradient, gather, integrator
# ---------------------------------------------------------------------------------------

def grad(phi: np.ndarray, dx: float):
    """g = -grad(phi), float32, reference deposit.py:132-174."""
    phi = phi.astype(F32)
    inv_dx = F32(1.0) / F32(dx)
    inv_2dx = F32(1.0) / (F32(2.0) * F32(dx))
    out = []
    for axis in range(3):
        g = np.empty_like(phi)
        p = np.moveaxis(phi, axis, 0)
        q = np.moveaxis(g, axis, 0)
        q[1:-1] = -(p[2:] - p[:-2]) * inv_2dx
        q[0] = -(p[1] - p[0]) * inv_dx
        q[-1] = -(p[-1] - p[-2]) * inv_dx
        out.append(g)
    return tuple(out)


def gather(pos: np.ndarray, gx: np.ndarray, gy: np.ndarray, gz: np.ndarray, dx: float):
    """CIC gather, float32, reference deposit.py:100-129 (corner order cx, cy, cz)."""
    n = gx.shape[0]
    axes = [_cic_axis(pos[:, a], dx, n) for a in range(3)]
    ok = axes[0][0] & axes[1][0] & axes[2][0]
    acc = [np.zeros(pos.shape[0], dtype=F32) for _ in range(3)]
    flat = [g.reshape(-1) for g in (gx, gy, gz)]
    for c in range(8):
        bits = (c >> 2, (c >> 1) & 1, c & 1)
        inb = ok.copy()
        ws = []
        idx = np.zeros(pos.shape[0], dtype=np.int64)
        for a, b in enumerate(bits):
            _, i, f = axes[a]
            j = i.astype(np.int64) + b
            inb &= (j >= 0) & (j < n)
            ws.append(f if b else (F32(1.0) - f).astype(F32))
            idx = idx * n + np.clip(j, 0, n - 1)
        w = (ws[0] * ws[1] * ws[2]).astype(F32)
        for q in range(3):
            term = (w * flat[q][idx]).astype(F32)
            acc[q] = np.where(inb, (acc[q] + term).astype(F32), acc[q])
    return tuple(acc)


def kick(v: np.ndarray, a: np.ndarray, dt: float) -> np.ndarray:
    return (v.astype(F32) + (a.astype(F32) * F32(dt)).astype(F32)).astype(F32)


def drift(x: np.ndarray, v: np.ndarray, dt: float) -> np.ndarray:
    return (x.astype(F32) + (v.astype(F32) * F32(dt)).astype(F32)).astype(F32)


def direct_accel_f64(pos: np.ndarray, mass: np.ndarray, grav: float, eps2: float) -> np.ndarray:
    """Float64 Plummer-softened direct sum (reference integrator.py:63-97)."""
    pos = pos.astype(np.float64)
    d = pos[None, :, :] - pos[:, None, :]
    r2 = (d ** 2).sum(-1) + eps2
    np.fill_diagonal(r2, np.inf)
    inv_r3 = r2 ** -1.5
    return grav * (mass[None, :, None] * d * inv_r3[:, :, None]).sum(1)


# ---------------------------------------------------------------------------------------
# multigrid pieces (float32, reference operation order)
# ---------------------------------------------------------------------------------------

def _interior(n):
    return (slice(1, n - 1),) * 3


def _sum6(p):
    n = p.shape[0]
    c = slice(1, n - 1)
    return (p[2:, c, c] + p[:-2, c, c] + p[c, 2:, c] + p[c, :-2, c] + p[c, c, 2:] + p[c, c, :-2]).astype(F32)


@functools.lru_cache(maxsize=32)
def _colour_mask(n: int, color: int) -> np.ndarray:
    ii, jj, kk = np.meshgrid(*(np.arange(1, n - 1),) * 3, indexing="ij")
    return ((ii + jj + kk) & 1) == color


def rb_pass(phi: np.ndarray, rhs: np.ndarray, h2: float, color: int) -> None:
    """In place: one colour of red-black Gauss-Seidel (multigrid.py:46-59)."""
    n = phi.shape[0]
    mask = _colour_mask(n, color)
    new = ((_sum6(phi) - (F32(h2) * rhs[_interior(n)]).astype(F32)) / F32(6.0)).astype(F32)
    inner = phi[_interior(n)]
    inner[mask] = new[mask]


def residual(phi: np.ndarray, rhs: np.ndarray, inv_h2: float) -> np.ndarray:
    """res = rhs - lap(phi) inside, 0 on the boundary (multigrid.py:62-79)."""
    n = phi.shape[0]
    res = np.zeros_like(phi, dtype=F32)
    lap = ((_sum6(phi) - (F32(6.0) * phi[_interior(n)]).astype(F32)) * F32(inv_h2)).astype(F32)
    res[_interior(n)] = rhs[_interior(n)] - lap
    return res


def restrict(res_f: np.ndarray) -> np.ndarray:
    """Full weighting onto the coarse grid of size nf//2 (multigrid.py:82-97)."""
    nf = res_f.shape[0]
    nc = nf // 2
    out = np.zeros((nc, nc, nc), dtype=F32)
    if nc < 3:
        return out
    acc = np.zeros((nc - 2,) * 3, dtype=F32)
    sl = lambda o: slice(2 + o, 2 * (nc - 1) + o, 2)  # fine 2J+o for J = 1..nc-2
    for a in (-1, 0, 1):
        for b in (-1, 0, 1):
            for c in (-1, 0, 1):
                w = F32((2.0 if a == 0 else 1.0) * (2.0 if b == 0 else 1.0) * (2.0 if c == 0 else 1.0))
                acc = (acc + (w * res_f[sl(a), sl(b), sl(c)]).astype(F32)).astype(F32)
    out[1:-1, 1:-1, 1:-1] = acc / F32(64.0)
    return out


def prolong_add(phi_c: np.ndarray, phi_f: np.ndarray) -> None:
    """In place: trilinear prolongation added into the fine interior (multigrid.py:100-130)."""
    nf = phi_f.shape[0]
    idx = np.arange(1, nf - 1)
    ci, oi = idx >> 1, idx & 1
    val = np.zeros((nf - 2,) * 3, dtype=F32)
    for di in (0, 1):
        wi = np.where(oi == 1, F32(0.5), F32(1.0) if di == 0 else F32(0.0)).astype(F32)
        for dj in (0, 1):
            wj = np.where(oi == 1, F32(0.5), F32(1.0) if dj == 0 else F32(0.0)).astype(F32)
            for dk in (0, 1):
                wk = np.where(oi == 1, F32(0.5), F32(1.0) if dk == 0 else F32(0.0)).astype(F32)
                w = (wi[:, None, None] * wj[None, :, None] * wk[None, None, :]).astype(F32)
                ai = np.minimum(ci + di, phi_c.shape[0] - 1)
                src = phi_c[np.ix_(ai, np.minimum(ci + dj, phi_c.shape[0] - 1),
                                   np.minimum(ci + dk, phi_c.shape[0] - 1))]
                term = (w * src).astype(F32)
                val = np.where(w > 0, (val + term).astype(F32), val)
    inner = phi_f[_interior(nf)]
    inner[...] = (inner + val).astype(F32)


def moments(rho: np.ndarray, dx: float) -> np.ndarray:
    """Ten raw float64 moments [M, Smx, Smy, Smz, Smxx, Smyy, Smzz, Smxy, Smxz, Smyz]."""
    n = rho.shape[0]
    m = rho.astype(np.float64) * float(dx) ** 3
    ax = np.arange(n, dtype=np.float64) * float(dx)
    mi, mj, mk = m.sum((1, 2)), m.sum((0, 2)), m.sum((0, 1))
    return np.array([
        m.sum(), (mi * ax).sum(), (mj * ax).sum(), (mk * ax).sum(),
        (mi * ax ** 2).sum(), (mj * ax ** 2).sum(), (mk * ax ** 2).sum(),
        ax @ m.sum(2) @ ax, ax @ m.sum(1) @ ax, ax @ m.sum(0) @ ax,
    ])


def face_params(s: np.ndarray, grav: float) -> np.ndarray:
    """[G, M, cx, cy, cz, sxx, syy, szz, sxy, sxz, syz] as float32 (multigrid.py:378-389)."""
    total = float(s[0])
    if total <= 0.0:
        vals = [0.0] * 10
    else:
        cx, cy, cz = s[1] / total, s[2] / total, s[3] / total
        vals = [total, cx, cy, cz,
                s[4] - total * cx * cx, s[5] - total * cy * cy, s[6] - total * cz * cz,
                s[7] - total * cx * cy, s[8] - total * cx * cz, s[9] - total * cy * cz]
    return np.array([grav] + list(vals), dtype=np.float64).astype(F32)


def face_value(i, j, k, dx: float, p: np.ndarray):
    """Multipole potential at nodes (i, j, k) (multigrid.py:279-295), float32."""
    g, mass, cmx, cmy, cmz, sxx, syy, szz, sxy, sxz, syz = (F32(v) for v in p)
    tr = F32(F32(sxx + syy) + szz)
    rx = (np.asarray(i).astype(F32) * F32(dx) - cmx).astype(F32)
    ry = (np.asarray(j).astype(F32) * F32(dx) - cmy).astype(F32)
    rz = (np.asarray(k).astype(F32) * F32(dx) - cmz).astype(F32)
    d2 = (rx * rx + ry * ry + rz * rz + F32(1e-12)).astype(F32)
    d = np.sqrt(d2).astype(F32)
    nx, ny, nz = rx / d, ry / d, rz / d
    nsn = (nx * nx * sxx + ny * ny * syy + nz * nz * szz
           + F32(2.0) * (nx * ny * sxy + nx * nz * sxz + ny * nz * syz)).astype(F32)
    quad = (F32(3.0) * nsn - tr).astype(F32)
    return (-g * (mass / d + quad / (F32(2.0) * d2 * d))).astype(F32)


def set_faces(phi: np.ndarray, dx: float, params: np.ndarray) -> None:
    n = phi.shape[0]
    ii, jj, kk = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    face = (ii == 0) | (ii == n - 1) | (jj == 0) | (jj == n - 1) | (kk == 0) | (kk == n - 1)
    phi[face] = face_value(ii[face], jj[face], kk[face], dx, params)


def level_sizes(n: int):
    sizes = [n]
    while sizes[-1] % 2 == 0 and sizes[-1] // 2 >= 3:
        sizes.append(sizes[-1] // 2)
    return sizes


def mg_solve(rho: np.ndarray, dx: float, grav: float, *, phi0: np.ndarray | None = None,
             n_cycles: int = 20, pre: int = 2, post: int = 2, coarse_sweeps: int = 40,
             tol: float = 1e-3, min_cycles: int = 1):
    """PLAIN multigrid solve (multigrid.py:455-487). Returns (phi, cycles, rhs_norm, resid)."""
    n = rho.shape[0]
    sizes = level_sizes(n)
    hs = [float(dx) * 2 ** l for l in range(len(sizes))]
    rhs = [None] * len(sizes)
    phi = [None] * len(sizes)
    rhs[0] = (F32(4.0 * math.pi * grav) * rho.astype(F32)).astype(F32)
    phi[0] = np.zeros((n, n, n), dtype=F32) if phi0 is None else phi0.astype(F32).copy()
    set_faces(phi[0], dx, face_params(moments(rho, dx), grav))

    def vcycle(l):
        h2 = F32(hs[l] ** 2)
        if l == len(sizes) - 1:
            for _ in range(coarse_sweeps):
                rb_pass(phi[l], rhs[l], h2, 0)
                rb_pass(phi[l], rhs[l], h2, 1)
            return
        for _ in range(pre):
            rb_pass(phi[l], rhs[l], h2, 0)
            rb_pass(phi[l], rhs[l], h2, 1)
        res = residual(phi[l], rhs[l], F32(1.0 / hs[l] ** 2))
        rhs[l + 1] = restrict(res)
        phi[l + 1] = np.zeros_like(rhs[l + 1])
        vcycle(l + 1)
        prolong_add(phi[l + 1], phi[l])
        for _ in range(post):
            rb_pass(phi[l], rhs[l], h2, 0)
            rb_pass(phi[l], rhs[l], h2, 1)

    rhs_norm = float(np.sqrt((rhs[0].astype(np.float64) ** 2).sum()))
    cycles = n_cycles
    resid = float("nan")
    for c in range(n_cycles):
        vcycle(0)
        if c + 1 >= min_cycles and rhs_norm > 0.0:
            r = residual(phi[0], rhs[0], F32(1.0 / hs[0] ** 2))
            resid = float(np.sqrt((r.astype(np.float64) ** 2).sum()))
            if resid < tol * rhs_norm:
                cycles = c + 1
                break
    return phi[0], cycles, rhs_norm, resid


# ---------------------------------------------------------------------------------------
# float64 ground truths for the Poisson problem
# ---------------------------------------------------------------------------------------

def dirichlet_solve_f64(f: np.ndarray, boundary: np.ndarray, dx: float) -> np.ndarray:
    """Exact float64 solution of lap7(phi) = f inside with phi = boundary on the faces.

    The interior system is diagonalised by the type-I discrete sine transform (a library
    routine, scipy.fft.dstn); boundary values are moved to the right-hand side first.
    """
    from scipy.fft import dstn, idstn

    n = f.shape[0]
    h2 = float(dx) ** 2
    b = (f.astype(np.float64) * h2)[_interior(n)].copy()
    bd = boundary.astype(np.float64).copy()
    bd[_interior(n)] = 0.0
    b -= (bd[2:, 1:-1, 1:-1] + bd[:-2, 1:-1, 1:-1] + bd[1:-1, 2:, 1:-1] + bd[1:-1, :-2, 1:-1]
          + bd[1:-1, 1:-1, 2:] + bd[1:-1, 1:-1, :-2])
    m = n - 2
    lam1 = 2.0 * (np.cos(np.pi * np.arange(1, m + 1) / (m + 1)) - 1.0)
    lam = lam1[:, None, None] + lam1[None, :, None] + lam1[None, None, :]
    u = idstn(dstn(b, type=1, workers=-1) / lam, type=1, workers=-1)
    out = boundary.astype(np.float64).copy()
    out[_interior(n)] = u
    return out


def fft_solve_f64(rho: np.ndarray, dx: float, grav: float, pad_factor: int = 2) -> np.ndarray:
    """Zero-padded continuum-kernel FFT solve (reference solver/fft_oracle.py:82-109)."""
    n = rho.shape[0]
    m = pad_factor * n
    d = np.minimum(np.arange(m), m - np.arange(m)).astype(np.float64)
    r2 = d[:, None, None] ** 2 + d[None, :, None] ** 2 + d[None, None, :] ** 2
    r2[0, 0, 0] = 1.0
    green = (-grav * dx ** 3 / dx) / np.sqrt(r2)
    green[0, 0, 0] = 0.0
    gft = np.fft.rfftn(green)
    padded = np.zeros((m, m, m))
    padded[:n, :n, :n] = rho.astype(np.float64)
    return np.fft.irfftn(np.fft.rfftn(padded) * gft, s=(m, m, m), axes=(0, 1, 2))[:n, :n, :n]


# ---------------------------------------------------------------------------------------
# diagnostics (float64)
# ---------------------------------------------------------------------------------------

def particle_scalars(pos, vel, mass) -> np.ndarray:
    """[M, KE, P(3), L(3) about the origin, Sum m r (3)] (diagnostics_device.py:41-74)."""
    m = mass.astype(np.float64)
    p = pos.astype(np.float64)
    v = vel.astype(np.float64)
    return np.concatenate([
        [math.fsum(m)], [math.fsum(0.5 * m * (v ** 2).sum(1))],
        (m[:, None] * v).sum(0), (m[:, None] * np.cross(p, v)).sum(0), (m[:, None] * p).sum(0)])


def grid_pe_raw(rho, phi) -> float:
    return math.fsum((rho.astype(np.float64) * phi.astype(np.float64)).ravel())


def lagrangian_radius(mass, pos, fraction=0.5) -> float:
    """Radius enclosing `fraction` of the mass about the CM (diagnostics.py:99-116)."""
    m = mass.astype(np.float64)
    p = pos.astype(np.float64)
    c = (m[:, None] * p).sum(0) / m.sum()
    r = np.linalg.norm(p - c, axis=1)
    order = np.argsort(r)
    cum = np.cumsum(m[order])
    idx = int(np.searchsorted(cum, fraction * m.sum()))
    return float(r[order][min(idx, len(r) - 1)])


# ---------------------------------------------------------------------------------------
# C port of the reference step (CPU baseline)
# ---------------------------------------------------------------------------------------

class CpuChain:
    """The reference PM step on host cores (OpenMP port in pm_oracle.c)."""

    def __init__(self, pos, vel, mass, n, dx, grav, threads: int | None = None):
        self.L = lib()
        if threads:
            self.L.cpu_set_threads(int(threads))
        self.threads = int(self.L.cpu_num_threads())
        c = lambda a: np.ascontiguousarray(a, dtype=F32)
        self.x, self.y, self.z = (c(pos[:, a]) for a in range(3))
        self.vx, self.vy, self.vz = (c(vel[:, a]) for a in range(3))
        self.m = c(mass)
        self.np = self.m.shape[0]
        self.n, self.dx = n, float(dx)
        shape = (n, n, n)
        self.rho, self.phi = np.zeros(shape, F32), np.zeros(shape, F32)
        self.gx, self.gy, self.gz = (np.zeros(shape, F32) for _ in range(3))
        self.ax, self.ay, self.az = (np.zeros(self.np, F32) for _ in range(3))
        self.mg = self.L.cpu_mg_create(n, self.dx, grav, 20, 2, 2, 40, 1, 1e-3)
        if not self.mg:
            raise MemoryError("cpu_mg_create failed")
        self.warm = 0
        self.last_cycles = 0

    # the stages of the reference step (deposit.py:47-174, multigrid solve, integrator.py:27-60)
    def deposit(self):
        self.L.cpu_deposit_atomic(_fp(self.x), _fp(self.y), _fp(self.z), _fp(self.m), self.np,
                                  self.n, self.dx, _fp(self.rho))

    def solve(self):
        self.last_cycles = self.L.cpu_mg_solve(self.mg, _fp(self.rho), _fp(self.phi), self.warm)
        self.warm = 1

    def grad(self):
        self.L.cpu_grad(_fp(self.phi), _fp(self.gx), _fp(self.gy), _fp(self.gz), self.n, self.dx)

    def gather(self):
        self.L.cpu_gather(_fp(self.x), _fp(self.y), _fp(self.z), _fp(self.gx), _fp(self.gy),
                          _fp(self.gz), _fp(self.ax), _fp(self.ay), _fp(self.az), self.np,
                          self.n, self.dx)

    def kick(self, h: float):
        self.L.cpu_kick(_fp(self.vx), _fp(self.vy), _fp(self.vz), _fp(self.ax), _fp(self.ay),
                        _fp(self.az), self.np, h)

    def drift(self, dt: float):
        self.L.cpu_drift(_fp(self.x), _fp(self.y), _fp(self.z), _fp(self.vx), _fp(self.vy),
                         _fp(self.vz), self.np, dt)

    def accel(self):
        self.deposit()
        self.solve()
        self.grad()
        self.gather()

    def step(self, dt: float):
        self.kick(0.5 * dt)
        self.drift(dt)
        self.accel()
        self.kick(0.5 * dt)

    def stage_times(self, dt: float) -> dict:
        """One step with each stage timed (seconds; the reference's per-stage protocol)."""
        import time

        out = {}

        def timed(name, fn, *a):
            t = time.perf_counter()
            fn(*a)
            out[name] = out.get(name, 0.0) + time.perf_counter() - t

        timed("integrate", self.kick, 0.5 * dt)
        timed("integrate", self.drift, dt)
        timed("deposit", self.deposit)
        timed("solve", self.solve)
        timed("grad", self.grad)
        timed("gather", self.gather)
        timed("integrate", self.kick, 0.5 * dt)
        return out

    def close(self):
        if self.mg:
            self.L.cpu_mg_destroy(self.mg)
            self.mg = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
