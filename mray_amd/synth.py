"""Synthetic inputs for the moment-image workloads (host / numpy side).

Every random choice comes from a counter-based SplitMix64 hash of
(seed, stream, counter), so a workload is a pure function of its seed. The
same formulas exist as HIP kernels (mray_amd/csrc/synth.hip) for the
full-size volumes; both sides are deterministic, but not bit-identical to each
other (libm vs device math), so parity tests always feed the same bytes to the
kernel and to the oracle.

Workloads (sizes as in the benchmark configurations):
  * Marschner-Lobb signal on [-1,1]^3 (fM = 6, a = 0.25), orthographic camera
    along +z;
  * mixing layer on [0,1]^3: 0.5 + 0.5 tanh((z - 0.5 - h(x,y)) / 0.01) plus
    4-octave value noise of amplitude 0.05, perspective camera;
  * large uint8 volume: 6-octave value noise through a smoothstep plus 64
    Gaussian blobs of 20-80 voxels (at 2048 x 2048 x 1920), seed 2;
  * 256-entry RGBA8 transfer functions made of piecewise-linear ramps;
  * SPH particles drawn from a mixture of anisotropic Gaussians.
"""
from __future__ import annotations

import math

import numpy as np

MASK64 = (1 << 64) - 1


def splitmix64(x):
    """SplitMix64 finaliser on uint64 numpy arrays (or ints)."""
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def hash_uniform(seed: int, stream: int, counter) -> np.ndarray:
    """Uniform doubles in [0, 1) from (seed, stream, counter)."""
    c = np.asarray(counter, dtype=np.uint64)
    with np.errstate(over="ignore"):
        key = splitmix64(np.uint64((seed * 0x100000001B3 + stream) & MASK64))
        h = splitmix64(key ^ splitmix64(c))
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def hash_normal(seed: int, stream: int, counter) -> np.ndarray:
    """Standard normals by Box-Muller over two hashed uniforms."""
    c = np.asarray(counter, dtype=np.uint64)
    u1 = hash_uniform(seed, stream, c * np.uint64(2))
    u2 = hash_uniform(seed, stream, c * np.uint64(2) + np.uint64(1))
    return np.sqrt(-2.0 * np.log(1.0 - u1)) * np.cos(2.0 * math.pi * u2)


# --------------------------------------------------------------------------
# volumes
# --------------------------------------------------------------------------

def voxel_coords(n: int, lo: float, hi: float) -> np.ndarray:
    """Voxel i sits at lo + i/n*(hi-lo) (RegularGrid::sample, src/volume.h:97-107)."""
    return lo + np.arange(n, dtype=np.float64) / n * (hi - lo)


def marschner_lobb(n: int, fm: float = 6.0, a: float = 0.25) -> np.ndarray:
    """Marschner-Lobb test signal sampled into an n^3 float32 grid over [-1,1]^3.

    f = (1 - sin(pi z / 2) + a (1 + cos(2 pi fM cos(pi r / 2)))) / (2 (1 + a)),
    r = sqrt(x^2 + y^2). Layout [z][y][x].
    """
    c = voxel_coords(n, -1.0, 1.0)
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    r = np.sqrt(x * x + y * y)
    rho_r = np.cos(2.0 * math.pi * fm * np.cos(math.pi * r / 2.0))
    f = (1.0 - np.sin(math.pi * z / 2.0) + a * (1.0 + rho_r)) / (2.0 * (1.0 + a))
    return np.ascontiguousarray(f.astype(np.float32))


def mixing_layer_modes(seed: int = 1, k_max: int = 16):
    """(A_k, theta_k, phi_k) for k = 1..k_max: A_k ~ N(0, 0.02/k), angles ~ U[0, 2pi)."""
    k = np.arange(1, k_max + 1)
    amp = hash_normal(seed, 1, k) * (0.02 / k)
    theta = 2.0 * math.pi * hash_uniform(seed, 2, k)
    phi = 2.0 * math.pi * hash_uniform(seed, 3, k)
    return amp, theta, phi


def _lattice_value(seed: int, octave: int, ix, iy, iz):
    """Value-noise lattice in [-1, 1] hashed from integer coordinates."""
    key = (np.asarray(ix, np.int64).astype(np.uint64) * np.uint64(73856093)) ^ \
          (np.asarray(iy, np.int64).astype(np.uint64) * np.uint64(19349663)) ^ \
          (np.asarray(iz, np.int64).astype(np.uint64) * np.uint64(83492791))
    return 2.0 * hash_uniform(seed, 100 + octave, key) - 1.0


def value_noise(x, y, z, seed: int, octaves: int, base: int) -> np.ndarray:
    """Sum of octaves with lattice base*2^o, weights 2^-o, normalised to [-1, 1]."""
    total = np.zeros(np.broadcast(x, y, z).shape)
    norm = 0.0
    for o in range(octaves):
        f = base * (1 << o)
        px, py, pz = x * f, y * f, z * f
        ix, iy, iz = np.floor(px), np.floor(py), np.floor(pz)
        fx, fy, fz = px - ix, py - iy, pz - iz
        sx, sy, sz = fx * fx * (3 - 2 * fx), fy * fy * (3 - 2 * fy), fz * fz * (3 - 2 * fz)
        acc = 0.0
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    w = (sx if dx else 1 - sx) * (sy if dy else 1 - sy) * (sz if dz else 1 - sz)
                    acc = acc + w * _lattice_value(seed, o, ix + dx, iy + dy, iz + dz)
        total = total + acc * (0.5 ** o)
        norm += 0.5 ** o
    return total / norm


def mixing_layer(n: int, seed: int = 1, width: float = 0.01) -> np.ndarray:
    """Mixing-layer density on [0,1]^3, n^3 float32, layout [z][y][x].

    width is the tanh thickness of the interface (0.01 in the benchmark)."""
    amp, theta, phi = mixing_layer_modes(seed)
    c = voxel_coords(n, 0.0, 1.0)
    y2, x2 = np.meshgrid(c, c, indexing="ij")
    h = np.zeros_like(x2)
    for k in range(16):
        h += amp[k] * np.sin(2 * math.pi * (k + 1) * (x2 * math.cos(theta[k]) + y2 * math.sin(theta[k])) + phi[k])
    out = np.empty((n, n, n), dtype=np.float32)
    for iz in range(n):
        z = c[iz]
        layer = 0.5 + 0.5 * np.tanh((z - 0.5 - h) / width)
        noise = value_noise(x2, y2, np.full_like(x2, z), seed, 4, 8)
        out[iz] = (layer + 0.05 * noise).astype(np.float32)
    return out


# the noise-and-blobs volume: normalised coordinates p = voxel index / nx on
# every axis (isotropic voxels, domain [0,1] x [0,ny/nx] x [0,nz/nx]);
# density v = clamp(BLOB_NOISE_WEIGHT * smoothstep(lo, hi, 0.5 + 0.5 noise6(p))
# + sum_k a_k exp(-|p - c_k|^2 / (2 s_k^2)), 0, 1), stored as floor(255 v + 0.5)
BLOB_FULL_SHAPE = (1920, 2048, 2048)  # (nz, ny, nx) of the full-size volume
BLOB_NOISE_BASE = 4                    # lattice of octave 0 (per unit of p)
BLOB_NOISE_OCTAVES = 6
BLOB_NOISE_WEIGHT = 0.5
BLOB_SMOOTH = (0.3, 0.7)               # smoothstep edges


def blob_params(seed: int = 2, n_blobs: int = 64):
    """Centres (n, 3) as (x, y, z) in normalised coordinates inside the
    full-size box, widths s_k from 20 to 80 voxels of the full-size volume
    (divided by its nx), amplitudes a_k from 0.4 to 1.0."""
    nz, ny, nx = BLOB_FULL_SHAPE
    b = np.arange(n_blobs)
    ext = np.array([1.0, ny / nx, nz / nx])
    centres = np.stack([hash_uniform(seed, 40 + d, b) * ext[d] for d in range(3)], axis=1)
    sigmas = (20.0 + 60.0 * hash_uniform(seed, 43, b)) / nx
    amps = 0.4 + 0.6 * hash_uniform(seed, 44, b)
    return centres.astype(np.float32), sigmas.astype(np.float32), amps.astype(np.float32)


def blob_volume_shape(scale: int = 1):
    """(nz, ny, nx) of the noise-and-blobs volume at 1/scale of full size."""
    return tuple(n // scale for n in BLOB_FULL_SHAPE)


# --------------------------------------------------------------------------
# transfer functions
# --------------------------------------------------------------------------

def tf_ramps(seed: int = 0, n_ramps: int = 4) -> np.ndarray:
    """256 x RGBA8 transfer function made of piecewise-linear ramps.

    Colour: linear interpolation between 6 hashed control colours at fixed
    positions. Opacity: n_ramps tent ramps with hashed centre, half-width and
    peak, summed and clipped; entry 0 stays transparent.
    """
    xs = np.arange(256, dtype=np.float64)
    ctrl_x = np.linspace(0, 255, 6)
    ctrl = hash_uniform(seed, 10, np.arange(18)).reshape(6, 3) * 255.0
    rgb = np.stack([np.interp(xs, ctrl_x, ctrl[:, c]) for c in range(3)], axis=1)
    alpha = np.zeros(256)
    centre = 30 + hash_uniform(seed, 11, np.arange(n_ramps)) * 200
    width = 8 + hash_uniform(seed, 12, np.arange(n_ramps)) * 40
    peak = 40 + hash_uniform(seed, 13, np.arange(n_ramps)) * 200
    for c, w, p in zip(centre, width, peak):
        alpha += np.clip(1.0 - np.abs(xs - c) / w, 0.0, None) * p
    alpha = np.clip(alpha, 0, 255)
    alpha[0] = 0
    tf = np.concatenate([rgb, alpha[:, None]], axis=1)
    return np.ascontiguousarray(np.floor(tf + 0.5).clip(0, 255).astype(np.uint8))


def tf_perturbed(tf_rgba8: np.ndarray, seed: int, sigma: float = 0.05) -> np.ndarray:
    """Float RGBA in [0,1]: tf/255 + N(0, sigma), clipped (transfer-function fitting start)."""
    noise = hash_normal(seed, 20, np.arange(1024)).reshape(256, 4) * sigma
    return np.clip(tf_rgba8.astype(np.float64) / 255.0 + noise, 0.0, 1.0)


# --------------------------------------------------------------------------
# SPH particles
# --------------------------------------------------------------------------

def particle_mixture(n: int, seed: int = 5, n_gauss: int = 32, box=(1.0, 1.0, 0.5)):
    """n particles from a mixture of anisotropic Gaussians inside box; uniform mass.

    Returns x, y, z (float32) and a per-particle value (float32, the
    normalised local density proxy 1.0, i.e. uniform mass).
    """
    box = np.asarray(box, dtype=np.float64)
    g = np.arange(n_gauss)
    centres = (0.15 + 0.7 * hash_uniform(seed, 30, g * 3 + np.arange(3)[:, None]).T) * box
    sig = (0.01 + 0.06 * hash_uniform(seed, 31, g * 3 + np.arange(3)[:, None]).T) * box
    idx = np.arange(n, dtype=np.uint64)
    comp = np.minimum((hash_uniform(seed, 32, idx) * n_gauss).astype(np.int64), n_gauss - 1)
    pts = np.empty((n, 3))
    for d in range(3):
        pts[:, d] = centres[comp, d] + sig[comp, d] * hash_normal(seed, 33 + d, idx)
    pts = np.clip(pts, 0.0, box * (1 - 1e-6))
    vals = np.ones(n, dtype=np.float32)
    return (pts[:, 0].astype(np.float32), pts[:, 1].astype(np.float32), pts[:, 2].astype(np.float32), vals)


# --------------------------------------------------------------------------
# cameras (mray_camera field order, see include/mray_types.h)
# --------------------------------------------------------------------------

def ortho_camera(pos=(0.0, 0.0, -3.0), direction=(0.0, 0.0, 1.0), right=(1.0, 0.0, 0.0), up=(0.0, 1.0, 0.0)):
    """Orthographic camera dict; znear is set explicitly to 0."""
    return dict(orthographic=1, transform=np.eye(4, dtype=np.float32).ravel().tolist(), position=list(pos),
                right=list(right), up=list(up), dir=list(direction), scale=1.0, aspect=1.0, znear=0.0)


def perspective_camera(position=(0.5, 0.5, -1.0), fov_half_deg: float = 25.0, znear: float = 0.01,
                       aspect: float = 1.0):
    """Perspective camera at `position` looking along world +z.

    The camera looks down its local -z axis; the camera-to-world transform is
    a rotation by pi about y (local -z -> world +z) plus the position in
    column 3, as cut::load_from_file expects (src/cut/camera.cpp:60-73).
    """
    px, py, pz = position
    m = np.array([[-1, 0, 0, px], [0, 1, 0, py], [0, 0, -1, pz], [0, 0, 0, 1]], dtype=np.float32)
    return dict(orthographic=0, transform=m.ravel().tolist(), position=[px, py, pz], right=[1, 0, 0],
                up=[0, 1, 0], dir=[0, 0, 1], scale=float(np.float32(math.tan(math.radians(fov_half_deg)))),
                aspect=aspect, znear=znear)
