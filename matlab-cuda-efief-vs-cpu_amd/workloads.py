"""Workload definitions of the EFIE patch sweep (used by bench.py and the tests).

Geometry family of the reference drivers (main_patch_efie_rwg.m:226-249):
a rectangular PEC patch at z=+h/2 over a PEC ground plate at z=-h/2, both
meshed by make_patch_with_ground_mesh.  Free space (eps = eps0).  The feed
point is on the patch edge at (-Wp/2, 0); the port edges are the nearest RWG
centers on each layer (pick_port_edge_by_point_and_layer.m).

Cell counts are square per layer because make_patch_with_ground_mesh.m
lists vertices y-fastest (meshgrid(...)(:), :5-7) but strides faces
x-fastest (:11-20); with Np_x != Np_y the face stride does not match the
vertex order and the patch layer overlaps itself.
"""
import numpy as np

C0 = 299792458.0
MU0 = 4 * np.pi * 1e-7
EPS0 = 1.0 / (MU0 * C0 ** 2)

# (Lp, Wp, Lg, Wg, h) in metres; Lp along y, Wp along x
GEOM = (30e-3, 38e-3, 60e-3, 60e-3, 1.6e-3)

CONFIGS = {
    # ~600 unknowns, 5 frequencies, the reference CPU path's size
    "c1": dict(geom=GEOM, cells=(8, 8, 12, 12), freqs=(2.0e9, 3.0e9, 5)),
    # N ~ 6k unknowns, 201 frequencies
    "c2_6k": dict(geom=GEOM, cells=(26, 26, 37, 37), freqs=(1.5e9, 3.5e9, 201)),
    # ~13k unknowns (the size class of 40x32 + 56x56 cells)
    "c2_13k": dict(geom=GEOM, cells=(36, 36, 56, 56), freqs=(1.5e9, 3.5e9, 201)),
    # tiny meshes for full-matrix parity
    "tiny": dict(geom=GEOM, cells=(3, 3, 4, 4), freqs=(2.0e9, 3.0e9, 5)),
    "micro": dict(geom=GEOM, cells=(2, 2, 3, 3), freqs=(2.0e9, 3.0e9, 3)),
}

# regular rule: 7-point; near/self: 6-point on 16 sub-triangles (accurate preset)
OPTS_C1 = dict(quad_order=7, near_sing_eps=2.5, self_subdiv=2, self_alpha=0.20,
               near_subdiv=2, near_alpha=0.15, near_order=4)


# Batched uneven sweep: geometries drawn from a seed (patch W, L uniform in
# 20-50 mm, h in 0.8-3.2 mm, ground 1.5-2.0x the patch, cell counts chosen so
# that N is uniform in 1000-8000 RWG), 64 frequencies in 1-4 GHz each.
C4 = dict(seed=20251203, n_geom=32, n_range=(1000, 8000), freqs=(1.0e9, 4.0e9, 64))


def c4_geometries(n_geom=None, n_range=None, seed=None):
    """[{name, geom (Lp, Wp, Lg, Wg, h), cells (np, np, ng, ng)}] of the
    batched uneven sweep.  Cell counts are square per layer (the mesh
    generator's stride, see the module docstring); with ng ~ f np the RWG
    count is N = 3 np^2 - 2 np + 3 ng^2 - 2 ng ~ 3 np^2 (1 + f^2)."""
    n_geom = C4["n_geom"] if n_geom is None else n_geom
    lo, hi = C4["n_range"] if n_range is None else n_range
    rng = np.random.Generator(np.random.PCG64(C4["seed"] if seed is None else seed))
    out = []
    for g in range(n_geom):
        Wp, Lp = rng.uniform(20e-3, 50e-3, 2)
        h = rng.uniform(0.8e-3, 3.2e-3)
        f = rng.uniform(1.5, 2.0)
        target = rng.uniform(lo, hi)
        npc = max(2, int(round(np.sqrt(target / (3.0 * (1.0 + f * f))))))
        ngc = max(npc + 1, int(round(f * npc)))
        out.append(dict(name=f"c4_{g}", geom=(Lp, Wp, f * Lp, f * Wp, h), cells=(npc, npc, ngc, ngc)))
    return out


def c4_feed_points(geom):
    Lp, Wp, Lg, Wg, h = geom
    return np.array([-Wp / 2, 0.0, +h / 2]), np.array([-Wp / 2, 0.0, -h / 2])


def freqs(name):
    a, b, n = CONFIGS[name]["freqs"]
    return np.linspace(a, b, n)


def feed_points(name):
    Lp, Wp, Lg, Wg, h = CONFIGS[name]["geom"]
    return np.array([-Wp / 2, 0.0, +h / 2]), np.array([-Wp / 2, 0.0, -h / 2])


def mesh_args(name):
    return (*CONFIGS[name]["geom"], *CONFIGS[name]["cells"])
