"""Porous-medium generators and the reference's geometry file formats.

Solid maps are ``int32`` arrays indexed ``[i][j][k]`` (x, y, z) with 0 = pore and 1 = solid,
the convention of the reference's geometry files (SP:1025-1065):

* binary: raw little-endian int32 in ``[i][j][k]`` C order (z fastest), ``decbin = 1``;
* ASCII:  whitespace-separated 0/1 tokens with x fastest, then y, then z, ``decbin = 0``;
  tokens other than 0 or 1 are skipped by the reference's reader (SP:1030-1034).
"""
from __future__ import annotations

import numpy as np


def write_binary(path: str, solid: np.ndarray) -> None:
    np.ascontiguousarray(solid, dtype="<i4").tofile(path)


def read_binary(path: str, shape) -> np.ndarray:
    a = np.fromfile(path, dtype="<i4")
    if a.size != int(np.prod(shape)):
        raise ValueError(f"{path}: {a.size} int32 values, expected {int(np.prod(shape))}")
    return a.reshape(shape).astype(np.int32)


def write_ascii(path: str, solid: np.ndarray) -> None:
    # x fastest: transpose [i][j][k] -> [k][j][i] and write in C order
    flat = np.ascontiguousarray(np.transpose(solid, (2, 1, 0))).reshape(-1)
    with open(path, "w") as fh:
        for start in range(0, flat.size, 64):
            fh.write(" ".join(str(int(v)) for v in flat[start:start + 64]))
            fh.write("\n")


def read_ascii(path: str, shape) -> np.ndarray:
    nx, ny, nz = shape
    vals = []
    need = nx * ny * nz
    with open(path) as fh:
        for line in fh:
            for tok in line.split():
                try:
                    v = float(tok)
                except ValueError:
                    continue
                if v == 0.0 or v == 1.0:
                    vals.append(int(v))
                    if len(vals) == need:
                        break
            if len(vals) == need:
                break
    if len(vals) != need:
        raise ValueError(f"{path}: {len(vals)} 0/1 tokens, expected {need}")
    return np.transpose(np.asarray(vals, np.int32).reshape(nz, ny, nx), (2, 1, 0)).copy()


def write_phase_ascii(path: str, psi: np.ndarray) -> None:
    """CG ``phase.dat``: floats, x fastest (CG:2079-2161)."""
    flat = np.ascontiguousarray(np.transpose(psi, (2, 1, 0))).reshape(-1)
    with open(path, "w") as fh:
        for start in range(0, flat.size, 32):
            fh.write(" ".join(repr(float(v)) for v in flat[start:start + 32]))
            fh.write("\n")


def sphere_pack(shape, radius: float, porosity: float, seed: int) -> np.ndarray:
    """Periodic box of randomly placed, overlapping solid spheres.

    Centres are drawn uniformly in the box from ``numpy.random.default_rng(seed)``; a voxel
    at integer position x is solid when its periodic minimum-image distance to a centre is
    at most ``radius``.  Spheres are added until the pore fraction is <= ``porosity``.
    """
    shape = tuple(int(s) for s in shape)
    rng = np.random.default_rng(seed)
    solid = np.zeros(shape, np.int32)
    total = solid.size
    n_solid = 0
    r = float(radius)
    ir = int(np.ceil(r))
    offs = np.arange(-ir - 1, ir + 2)
    while (total - n_solid) / total > porosity:
        c = rng.uniform(0.0, 1.0, 3) * np.asarray(shape)
        base = np.floor(c).astype(int)
        sel = []
        for a in range(3):
            coords = base[a] + offs
            d = coords - c[a]
            sel.append((np.mod(coords, shape[a]), d))
        (ix, dx), (iy, dy), (iz, dz) = sel
        d2 = dx[:, None, None] ** 2 + dy[None, :, None] ** 2 + dz[None, None, :] ** 2
        inside = d2 <= r * r
        sub = solid[np.ix_(ix, iy, iz)]
        newly = inside & (sub == 0)
        n_new = int(np.count_nonzero(newly))
        if n_new:
            sub[newly] = 1
            solid[np.ix_(ix, iy, iz)] = sub
            n_solid += n_new
    return solid


def gaussian_random_field(shape, corr_len: float, seed: int) -> np.ndarray:
    """Periodic Gaussian random field: white noise smoothed by a Gaussian kernel of std
    ``corr_len`` (applied in Fourier space)."""
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal(shape)
    k = [np.fft.fftfreq(n) * 2 * np.pi for n in shape]
    k2 = k[0][:, None, None] ** 2 + k[1][None, :, None] ** 2 + k[2][None, None, :] ** 2
    spec = np.fft.fftn(noise) * np.exp(-0.5 * k2 * corr_len ** 2)
    field = np.real(np.fft.ifftn(spec))
    return (field - field.mean()) / (field.std() + 1e-300)


def grf_sample(edge: int, corr_len: float, porosity: float, seed: int) -> np.ndarray:
    """Thresholded Gaussian random field: the lowest ``porosity`` quantile is pore."""
    field = gaussian_random_field((edge, edge, edge), corr_len, seed)
    thr = np.quantile(field, porosity)
    return (field > thr).astype(np.int32)


def fibre_pack(shape, radius: float, porosity: float, seed: int) -> np.ndarray:
    """Periodic box of straight solid fibres (cylinders of ``radius``) with random centres and
    isotropic orientations from ``numpy.random.default_rng(seed)``, added until the pore
    fraction is <= ``porosity``.

    A fibre with centre c and unit direction v spans one box length along the dominant axis a
    of v: in every plane x_a = k its axis passes through p_k = c + v d_k / v_a, with d_k the
    periodic minimum image of k - c_a, and a voxel y of that plane is solid when its distance
    to the axis line, |y - p_k|^2 - ((y - p_k).v)^2, is at most radius^2.  Indices wrap
    periodically.  The in-plane cross-section is an ellipse with semi-axes radius and
    radius/|v_a| <= radius*sqrt(3), so each plane needs only a small window of candidates.
    """
    shape = tuple(int(s) for s in shape)
    L = np.asarray(shape, np.float64)
    rng = np.random.default_rng(seed)
    solid = np.zeros(shape, bool)
    flat = solid.reshape(-1)
    total = solid.size
    n_solid = 0
    r2 = float(radius) ** 2
    strides = np.array([shape[1] * shape[2], shape[2], 1], np.int64)
    while (total - n_solid) / total > porosity:
        c = rng.uniform(0, 1, 3) * L
        v = rng.standard_normal(3)
        v /= np.linalg.norm(v)
        a = int(np.argmax(np.abs(v)))
        b, e = [x for x in range(3) if x != a]
        h = int(np.ceil(radius / abs(v[a]))) + 1
        if 2 * h + 1 > min(shape[b], shape[e]):
            raise ValueError("fibre radius too large for the box")
        k = np.arange(shape[a])
        dk = k - c[a]
        dk -= L[a] * np.round(dk / L[a])
        p = c[None, :] + v[None, :] * (dk / v[a])[:, None]               # axis point per plane
        off = np.arange(-h, h + 1)
        yb = np.floor(p[:, b])[:, None, None] + off[None, :, None]        # (La, W, 1)
        ye = np.floor(p[:, e])[:, None, None] + off[None, None, :]        # (La, 1, W)
        db = yb - p[:, b, None, None]
        de = ye - p[:, e, None, None]
        t = db * v[b] + de * v[e]                                          # y - p_k has no a part
        inside = db * db + de * de - t * t <= r2
        kk = np.broadcast_to(k[:, None, None], inside.shape)[inside]
        ib = np.mod(np.broadcast_to(yb, inside.shape)[inside].astype(np.int64), shape[b])
        ie = np.mod(np.broadcast_to(ye, inside.shape)[inside].astype(np.int64), shape[e])
        lin = kk * strides[a] + ib * strides[b] + ie * strides[e]
        n_solid += int(np.count_nonzero(~flat[lin]))
        flat[lin] = True
    return solid.astype(np.int32)


def pores_wrap(solid: np.ndarray, axis: int) -> bool:
    """Whether a cluster of pore voxels, joined by the 18 D3Q19 links with the box periodic on
    every axis, wraps round the box along ``axis`` (0 = x, 1 = y, 2 = z).  Without one there is
    no throughflow along the axis at steady state.  Host code of the library (no device use):
    ``lbm_geometry_wraps`` of the C ABI."""
    import ctypes
    from . import _capi
    s = np.ascontiguousarray(solid, dtype=np.int32)
    if s.ndim != 3:
        raise ValueError("solid must be a 3-D [i][j][k] array")
    out = ctypes.c_int32(0)
    rc = _capi.lib().lbm_geometry_wraps(int(s.shape[0]), int(s.shape[1]), int(s.shape[2]),
                                        s.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), int(axis),
                                        ctypes.byref(out))
    _capi.check(rc, None)
    return bool(out.value)


def porosity(solid: np.ndarray) -> float:
    return float(np.count_nonzero(solid == 0)) / solid.size
