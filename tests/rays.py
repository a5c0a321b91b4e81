"""Seeded ray sets for the traversal parity tests (8 floats per ray:
o.xyz, tmax, d.xyz, pad)."""
import numpy as np


def pack(o, d, tmax=1e16):
    o = np.asarray(o, np.float32).reshape(-1, 3)
    d = np.asarray(d, np.float32).reshape(-1, 3)
    r = np.zeros((len(o), 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = tmax
    r[:, 4:7] = d
    return r


def random_rays(n, lo, hi, seed=0):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return pack(o, d)


def adversarial_cornell(seed=0):
    """Rays that stress ties and slab edge cases in the Cornell room:
    axis-parallel directions (zero components), origins on grid lines and
    vertices shared by several triangles, rays grazing along the walls,
    origins inside the boxes."""
    rng = np.random.default_rng(seed)
    rays = []
    axes = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.float32)
    grid = np.linspace(-1, 1, 17)
    # axis-parallel rays from grid-aligned origins: hit shared edges/vertices
    for _ in range(400):
        o = np.array([rng.choice(grid), rng.choice(grid) + 1.0, rng.choice(grid)])
        o = np.clip(o, [-0.999, 0.001, -0.999], [0.999, 1.999, 0.999])
        rays.append(pack(o, axes[rng.integers(6)]))
    # rays aimed exactly at wall grid vertices
    for _ in range(400):
        target = np.array([rng.choice(grid), 0.0, rng.choice(grid)])
        o = rng.uniform([-0.8, 0.2, -0.8], [0.8, 1.8, 0.8])
        d = target - o
        rays.append(pack(o, d / np.linalg.norm(d)))
    # grazing: direction in the plane of a wall, origin on the wall plane
    for _ in range(200):
        o = np.array([rng.uniform(-0.9, 0.9), 0.0, rng.uniform(-0.9, 0.9)])
        phi = rng.uniform(0, 2 * np.pi)
        rays.append(pack(o, [np.cos(phi), 0.0, np.sin(phi)]))
    # origins inside the boxes / sphere
    for _ in range(200):
        o = rng.uniform([-0.65, 0.05, -0.6], [-0.15, 1.15, -0.1])
        d = rng.normal(size=3)
        rays.append(pack(o, d / np.linalg.norm(d)))
    # short segments (shadow-ray style tmax)
    r = random_rays(300, [-0.9, 0.1, -0.9], [0.9, 1.9, 0.9], seed + 7)
    r[:, 3] = rng.uniform(0.01, 1.0, len(r))
    rays.append(r)
    return np.concatenate(rays).astype(np.float32)
