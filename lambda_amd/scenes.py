"""Generators for the benchmark scenes, written as .lrd (+ .obj) files.

Every scene is deterministic: positions that the workloads describe as
random come from numpy's PCG64 generator with a fixed seed and this module's
own uniform mapping.  The files use only the .lrd grammar of Lambda's parser
(README.md:20-82, core/parser.h:62-324).

Encoding choices forced by the reference's semantics:

* Dielectric ``n0``/``n1`` are the outside/inside refractive-index spectra
  (scene/material.h:185-226, core/parser.h:171-179), so a Cauchy glass
  n(lambda) = A + B / lambda_um^2 is written as ``n0 = 1`` and ``n1`` a
  sampled spectrum.
* Checker cells are in (u, v) units (scene/texture.h:203-208); a quad has
  u, v in [0, 1], so "scale 10" (ten cells per side) is written as 0.1.
* A scalar texture is used as a spectrum through the implicit wrapping of
  parseSpectrumTexture (core/parser.h:628-633) rather than the
  ``spectrum scalar`` texture command, which reads its name twice
  (core/parser.h:141).
"""
from __future__ import annotations

import math
import os

import numpy as np

__all__ = [
    "write_closed_box",
    "write_thinfilm_grid",
    "write_instanced_mesh",
    "write_guiding_room",
    "icosphere",
    "perlin3",
]


def _f(x: float) -> str:
    return "%.9g" % float(np.float32(x))


def _v(v) -> str:
    return "(%s %s %s)" % (_f(v[0]), _f(v[1]), _f(v[2]))


def look_at_camera(position, target, up, vfov_deg, aspect, plane_distance=1.0):
    """Camera command fields (position, corner, horizontal, vertical).

    ``corner`` is the top-left of the image plane; ``vertical`` points down
    the image so that row 0 is the top (render/camera.h:13-16).
    """
    p = np.asarray(position, np.float64)
    fwd = np.asarray(target, np.float64) - p
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, np.asarray(up, np.float64))
    right /= np.linalg.norm(right)
    true_up = np.cross(right, fwd)
    half_h = math.tan(math.radians(vfov_deg) / 2) * plane_distance
    half_w = half_h * aspect
    centre = p + fwd * plane_distance
    corner = centre - right * half_w + true_up * half_h
    horizontal = right * (2 * half_w)
    vertical = -true_up * (2 * half_h)
    return p, corner, horizontal, vertical


def _camera_line(position, corner, horizontal, vertical) -> str:
    return "Camera %s %s %s %s" % (_v(position), _v(corner), _v(horizontal), _v(vertical))


def _cauchy_spectrum(a: float, b_um2: float, lo: int = 360, hi: int = 830, step: int = 10) -> str:
    pts = []
    for lam in range(lo, hi + 1, step):
        n = a + b_um2 / (lam / 1000.0) ** 2
        pts.append("%d %s" % (lam, _f(n)))
    return "( " + " ".join(pts) + " )"


# --------------------------------------------------------------------------
# Closed box: 5 Lambertian walls, a ceiling emitter, a mirror and a glass ball.


def write_closed_box(directory: str, width=256, height=256, depth=8, samples=16,
                     lambda_min=380, lambda_max=780, name="closed_box") -> str:
    """Small closed box (all geometry in [0,1]^3), camera at (0.5,0.5,-1.4)."""
    os.makedirs(directory, exist_ok=True)
    lines = [
        "# Closed box: 5 wall quads, ceiling emitter, mirror and dielectric spheres",
        "Render %d %d %d %d %s %s" % (width, height, depth, samples, _f(lambda_min), _f(lambda_max)),
        "Material white lambertian 0.75",
        "Material red lambertian ( 599 0 600 0.65 )",
        "Material green lambertian ( 499 0 500 0.6 570 0.6 571 0 )",
        "Material light emissive 15",
        "Material chrome mirror 0.9",
        # Cauchy glass n = 1.5 + 0.0042 / lambda_um^2 inside, air outside.
        "Material glass dielectric 1 %s" % _cauchy_spectrum(1.5, 0.0042),
        "Object box quad white (0 0 0) (1 0 0) (0 0 1)",      # floor  (y = 0)
        "Object box quad white (0 1 0) (1 0 0) (0 0 1)",      # ceiling (y = 1)
        "Object box quad white (0 0 1) (1 0 0) (0 1 0)",      # back   (z = 1)
        "Object box quad red (0 0 0) (0 1 0) (0 0 1)",        # x = 0
        "Object box quad green (1 0 0) (0 1 0) (0 0 1)",      # x = 1
        "Object lamp quad light (0.375 0.999 0.375) (0.25 0 0) (0 0 0.25)",
        "Object balls sphere chrome (0.3 0.18 0.62) 0.18",
        "Object balls sphere glass (0.7 0.18 0.38) 0.18",
        "Instance box default (0 0 0) (0 0 0) (1 1 1)",
        "Instance lamp default (0 0 0) (0 0 0) (1 1 1)",
        "Instance balls default (0 0 0) (0 0 0) (1 1 1)",
        "Background equirectangular 0",
    ]
    cam = look_at_camera((0.5, 0.5, -1.4), (0.5, 0.5, 0.5), (0, 1, 0), 38.0, width / height)
    lines.append(_camera_line(*cam))
    path = os.path.join(directory, name + ".lrd")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return path


# --------------------------------------------------------------------------
# Wave-optical materials: thin-film spheres on a checker floor.


def write_thinfilm_grid(directory: str, width=512, height=512, depth=12, samples=64,
                        lambda_min=360, lambda_max=830, seed=7, name="thinfilm_grid") -> str:
    """3x3 grid of r=0.3 soap-film spheres (n=1.33 in air) with perlin or
    worley thickness, checker floor, perlin environment."""
    os.makedirs(directory, exist_ok=True)
    rng = np.random.Generator(np.random.PCG64(seed))
    jitter = rng.random((9, 2)) * 0.1 - 0.05
    lines = [
        "# Thin-film spheres over a checker floor under a perlin environment",
        "Render %d %d %d %d %s %s" % (width, height, depth, samples, _f(lambda_min), _f(lambda_max)),
        "Texture perlinD scalar perlin 200 800 4",
        "Texture worleyD scalar worley 100 600 8",
        "Texture sky scalar perlin 0 1 2",
        "Texture floorTex spectrum checker 0.8 0.1 0.1",
        "Material filmP thinfilm [ 1 1.33 1 ] [ perlinD ]",
        "Material filmW thinfilm [ 1 1.33 1 ] [ worleyD ]",
        "Material floor lambertian floorTex",
        "Object floor quad floor (-2.5 0 -2.5) (5 0 0) (0 0 5)",
    ]
    k = 0
    for i in range(3):
        for j in range(3):
            x = (i - 1) * 0.8 + jitter[k, 0]
            z = (j - 1) * 0.8 + jitter[k, 1]
            mat = "filmW" if k in (1, 4, 8) else "filmP"
            lines.append("Object spheres sphere %s (%s 0.3 %s) 0.3" % (mat, _f(x), _f(z)))
            k += 1
    lines += [
        "Instance floor default (0 0 0) (0 0 0) (1 1 1)",
        "Instance spheres default (0 0 0) (0 0 0) (1 1 1)",
        "Background equirectangular sky",
    ]
    cam = look_at_camera((0.0, 1.6, -3.2), (0.0, 0.25, 0.0), (0, 1, 0), 40.0, width / height)
    lines.append(_camera_line(*cam))
    path = os.path.join(directory, name + ".lrd")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return path


# --------------------------------------------------------------------------
# Improved-noise Perlin (Perlin 2002) in numpy, for mesh displacement.

_PERM = np.array([
    151, 160, 137, 91, 90, 15, 131, 13, 201, 95, 96, 53, 194, 233, 7, 225, 140, 36, 103, 30, 69, 142,
    8, 99, 37, 240, 21, 10, 23, 190, 6, 148, 247, 120, 234, 75, 0, 26, 197, 62, 94, 252, 219, 203, 117,
    35, 11, 32, 57, 177, 33, 88, 237, 149, 56, 87, 174, 20, 125, 136, 171, 168, 68, 175, 74, 165, 71,
    134, 139, 48, 27, 166, 77, 146, 158, 231, 83, 111, 229, 122, 60, 211, 133, 230, 220, 105, 92, 41,
    55, 46, 245, 40, 244, 102, 143, 54, 65, 25, 63, 161, 1, 216, 80, 73, 209, 76, 132, 187, 208, 89,
    18, 169, 200, 196, 135, 130, 116, 188, 159, 86, 164, 100, 109, 198, 173, 186, 3, 64, 52, 217, 226,
    250, 124, 123, 5, 202, 38, 147, 118, 126, 255, 82, 85, 212, 207, 206, 59, 227, 47, 16, 58, 17, 182,
    189, 28, 42, 223, 183, 170, 213, 119, 248, 152, 2, 44, 154, 163, 70, 221, 153, 101, 155, 167, 43,
    172, 9, 129, 22, 39, 253, 19, 98, 108, 110, 79, 113, 224, 232, 178, 185, 112, 104, 218, 246, 97,
    228, 251, 34, 242, 193, 238, 210, 144, 12, 191, 179, 162, 241, 81, 51, 145, 235, 249, 14, 239, 107,
    49, 192, 214, 31, 181, 199, 106, 157, 184, 84, 204, 176, 115, 121, 50, 45, 127, 4, 150, 254, 138,
    236, 205, 93, 222, 114, 67, 29, 24, 72, 243, 141, 128, 195, 78, 66, 215, 61, 156, 180], np.int64)


def _grad(h, x, y, z):
    h = h % 12
    u = np.where(h < 8, x, y)
    v = np.where(h < 4, y, np.where(h < 8, z, z))
    # 12 edge directions: (+-x +-y), (+-x +-z), (+-y +-z)
    a = np.where((h & 1) == 0, 1.0, -1.0)
    b = np.where((h & 2) == 0, 1.0, -1.0)
    return a * u + b * v


def perlin3(p: np.ndarray) -> np.ndarray:
    """Improved Perlin noise remapped to [0, 1] for points p[..., 3]."""
    p = np.asarray(p, np.float64)
    fl = np.floor(p)
    i = fl.astype(np.int64) & 255
    t = p - fl
    f = t * t * t * (t * (t * 6 - 15) + 10)
    P = lambda k: _PERM[k & 255]  # noqa: E731
    xi, yi, zi = i[..., 0], i[..., 1], i[..., 2]
    tx, ty, tz = t[..., 0], t[..., 1], t[..., 2]

    def corner(dx, dy, dz):
        h = P(P(P(xi + dx) + yi + dy) + zi + dz)
        return _grad(h, tx - dx, ty - dy, tz - dz)

    lerp = lambda a, b, w: a + w * (b - a)  # noqa: E731
    x00 = lerp(corner(0, 0, 0), corner(1, 0, 0), f[..., 0])
    x10 = lerp(corner(0, 1, 0), corner(1, 1, 0), f[..., 0])
    x01 = lerp(corner(0, 0, 1), corner(1, 0, 1), f[..., 0])
    x11 = lerp(corner(0, 1, 1), corner(1, 1, 1), f[..., 0])
    y0 = lerp(x00, x10, f[..., 1])
    y1 = lerp(x01, x11, f[..., 1])
    return (lerp(y0, y1, f[..., 2]) + 1) / 2


def icosphere(level: int):
    """Unit icosphere: 20 * 4**level triangles. Returns (vertices, faces)."""
    t = (1 + 5 ** 0.5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
         (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = np.array(v, np.float64)
    verts /= np.linalg.norm(verts, axis=1, keepdims=True)
    faces = np.array(f, np.int64)
    for _ in range(level):
        nv = len(verts)
        e = np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]])
        e.sort(axis=1)
        key = e[:, 0] * (nv + 1) + e[:, 1]
        uniq, inv = np.unique(key, return_inverse=True)
        a = uniq // (nv + 1)
        b = uniq % (nv + 1)
        mids = verts[a] + verts[b]
        mids /= np.linalg.norm(mids, axis=1, keepdims=True)
        verts = np.concatenate([verts, mids])
        m = inv.reshape(3, -1) + nv
        m01, m12, m20 = m[0], m[1], m[2]
        f0, f1, f2 = faces[:, 0], faces[:, 1], faces[:, 2]
        faces = np.concatenate([
            np.stack([f0, m01, m20], 1), np.stack([f1, m12, m01], 1),
            np.stack([f2, m20, m12], 1), np.stack([m01, m12, m20], 1)])
    return verts, faces


def write_obj(path: str, verts: np.ndarray, faces: np.ndarray) -> None:
    v32 = verts.astype(np.float32)
    with open(path, "w") as fh:
        fh.write("# generated mesh: %d vertices, %d triangles\n" % (len(v32), len(faces)))
        np.savetxt(fh, v32, fmt="v %.9g %.9g %.9g")
        np.savetxt(fh, faces + 1, fmt="f %d %d %d")


def write_instanced_mesh(directory: str, width=1920, height=1080, depth=10, samples=64,
                         level=8, grid=8, amplitude=0.05, frequency=6.0, seed=11,
                         lambda_min=380, lambda_max=780, name="instanced_mesh") -> str:
    """Perlin-displaced icosphere instanced grid x grid times with seeded
    rotation and uniform scale 0.5-1.5; materials round-robin Lambertian,
    mirror, dielectric n=1.5, thin film d=400 nm; 4 quad emitters plus a
    constant environment."""
    os.makedirs(directory, exist_ok=True)
    obj = os.path.join(directory, "%s_ico%d.obj" % (name, level))
    if not os.path.exists(obj):
        verts, faces = icosphere(level)
        r = 1.0 + amplitude * (2.0 * perlin3(verts * frequency) - 1.0)
        write_obj(obj + ".tmp", verts * r[:, None], faces)
        os.replace(obj + ".tmp", obj)
    rng = np.random.Generator(np.random.PCG64(seed))
    rot = rng.random((grid * grid, 3)) * (2 * math.pi)
    scl = 0.5 + rng.random(grid * grid)
    spacing = 3.2
    half = (grid - 1) * spacing / 2
    lines = [
        "# %dx%d instances of a level-%d perlin-displaced icosphere" % (grid, grid, level),
        "Render %d %d %d %d %s %s" % (width, height, depth, samples, _f(lambda_min), _f(lambda_max)),
        "Material matte lambertian 0.7",
        "Material chrome mirror 0.9",
        "Material glass dielectric 1 1.5",
        "Material film thinfilm [ 1 1.33 1.5 ] [ 400 ]",
        "Material light emissive 8",
        "Object mesh mesh matte %s" % obj,
    ]
    extent = half + spacing
    for k, (x, z) in enumerate([(-1, -1), (1, -1), (-1, 1), (1, 1)]):
        cx, cz = x * extent * 0.5, z * extent * 0.5
        lines.append("Object lamps quad light (%s 9 %s) (4 0 0) (0 0 4)" % (_f(cx - 2), _f(cz - 2)))
    mats = ["matte", "chrome", "glass", "film"]
    for k in range(grid * grid):
        gx, gz = k % grid, k // grid
        t = (gx * spacing - half, 1.6, gz * spacing - half)
        lines.append("Instance mesh %s %s %s (%s %s %s)" % (
            mats[k % 4], _v(t), _v(rot[k]), _f(scl[k]), _f(scl[k]), _f(scl[k])))
    lines += [
        "Instance lamps default (0 0 0) (0 0 0) (1 1 1)",
        "Background equirectangular 0.05",
    ]
    cam = look_at_camera((0.0, 18.0, -30.0), (0.0, 0.0, 0.0), (0, 1, 0), 42.0, width / height)
    lines.append(_camera_line(*cam))
    path = os.path.join(directory, name + ".lrd")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return path


# --------------------------------------------------------------------------
# Hard lighting: a closed room lit only through a narrow door gap.


def _cube_quads(c, s):
    x, y, z = c
    h = s / 2
    x0, y0, z0 = x - h, y - h, z - h
    return [
        ((x0, y0, z0), (s, 0, 0), (0, 0, s)),      # bottom
        ((x0, y0 + s, z0), (s, 0, 0), (0, 0, s)),  # top
        ((x0, y0, z0), (s, 0, 0), (0, s, 0)),      # front
        ((x0, y0, z0 + s), (s, 0, 0), (0, s, 0)),  # back
        ((x0, y0, z0), (0, s, 0), (0, 0, s)),      # left
        ((x0 + s, y0, z0), (0, s, 0), (0, 0, s)),  # right
    ]


def write_guiding_room(directory: str, width=1920, height=1080, depth=16, samples=1, clutter=200,
                       seed=5, lambda_min=380, lambda_max=780, name="guiding_room") -> str:
    """Closed 4x3x4 room; a 0.1x0.1 emitter (500) sits in an adjoining box
    reachable only through a 0.05-wide gap in the shared wall; 200 seeded
    Lambertian spheres and cubes."""
    os.makedirs(directory, exist_ok=True)
    rng = np.random.Generator(np.random.PCG64(seed))
    W, H, D = 4.0, 3.0, 4.0
    gap = 0.05
    lines = [
        "# Room lit through a %.2f-wide door gap from an adjoining box" % gap,
        "Render %d %d %d %d %s %s" % (width, height, depth, samples, _f(lambda_min), _f(lambda_max)),
        "Material wall lambertian 0.7",
        "Material clutterA lambertian 0.6",
        "Material clutterB lambertian ( 380 0.2 550 0.7 780 0.4 )",
        "Material light emissive 500",
        # room: floor, ceiling, front (z=0), left (x=0); the right wall x=W has a gap
        "Object room quad wall (0 0 0) (%s 0 0) (0 0 %s)" % (_f(W), _f(D)),
        "Object room quad wall (0 %s 0) (%s 0 0) (0 0 %s)" % (_f(H), _f(W), _f(D)),
        "Object room quad wall (0 0 0) (%s 0 0) (0 %s 0)" % (_f(W), _f(H)),
        "Object room quad wall (0 0 %s) (%s 0 0) (0 %s 0)" % (_f(D), _f(W), _f(H)),
        "Object room quad wall (0 0 0) (0 %s 0) (0 0 %s)" % (_f(H), _f(D)),
    ]
    zc = D / 2
    # right wall x=W split around a vertical gap [zc - gap/2, zc + gap/2]
    lines.append("Object room quad wall (%s 0 0) (0 %s 0) (0 0 %s)" % (_f(W), _f(H), _f(zc - gap / 2)))
    lines.append("Object room quad wall (%s 0 %s) (0 %s 0) (0 0 %s)" % (_f(W), _f(zc + gap / 2), _f(H), _f(D - zc - gap / 2)))
    # adjoining box x in [W, W+1], z in [zc-0.5, zc+0.5]
    bx = W + 1.0
    lines += [
        "Object annex quad wall (%s 0 %s) (1 0 0) (0 0 1)" % (_f(W), _f(zc - 0.5)),
        "Object annex quad wall (%s %s %s) (1 0 0) (0 0 1)" % (_f(W), _f(H), _f(zc - 0.5)),
        "Object annex quad wall (%s 0 %s) (1 0 0) (0 %s 0)" % (_f(W), _f(zc - 0.5), _f(H)),
        "Object annex quad wall (%s 0 %s) (1 0 0) (0 %s 0)" % (_f(W), _f(zc + 0.5), _f(H)),
        "Object annex quad wall (%s 0 %s) (0 %s 0) (0 0 1)" % (_f(bx), _f(zc - 0.5), _f(H)),
        "Object lamp quad light (%s %s %s) (0.1 0 0) (0 0 0.1)" % (_f(bx - 0.3), _f(H - 0.01), _f(zc - 0.05)),
    ]
    for k in range(clutter):
        mat = "clutterA" if k % 2 == 0 else "clutterB"
        s = 0.08 + 0.17 * rng.random()
        c = (0.3 + (W - 0.6) * rng.random(), s / 2 + 0.001 + 1.2 * rng.random(), 0.3 + (D - 0.6) * rng.random())
        if k % 2 == 0:
            lines.append("Object clutter sphere %s %s %s" % (mat, _v(c), _f(s / 2)))
        else:
            for corner, hh, vv in _cube_quads(c, s):
                lines.append("Object clutter quad %s %s %s %s" % (mat, _v(corner), _v(hh), _v(vv)))
    lines += [
        "Instance room default (0 0 0) (0 0 0) (1 1 1)",
        "Instance annex default (0 0 0) (0 0 0) (1 1 1)",
        "Instance lamp default (0 0 0) (0 0 0) (1 1 1)",
        "Instance clutter default (0 0 0) (0 0 0) (1 1 1)",
        "Background equirectangular 0",
    ]
    cam = look_at_camera((0.4, 1.6, 0.3), (3.5, 1.0, 3.0), (0, 1, 0), 60.0, width / height, 0.2)
    lines.append(_camera_line(*cam))
    path = os.path.join(directory, name + ".lrd")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return path
