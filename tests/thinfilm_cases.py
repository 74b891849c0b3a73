"""Thin-film stacks and query sets shared by the CPU and GPU thin-film tests."""
import numpy as np

import lambda_amd

# (indices outer..substrate, thicknesses in nm): lossless single films, a bare
# interface, AR multilayers, an absorbing film, metals, a dense-to-rare
# stack (total internal reflection from the back).
STACKS = [
    ([1, 1.33, 1], [350.0]),
    ([1, 1.5], []),
    ([1, 1.38, 2.1, 1.52], [99.6, 71.4]),
    ([1, 1.33 + 0.02j, 1.5], [600.0]),
    ([1, 0.2 + 3.4j], []),
    ([1, 2.0, 1.4 + 0.3j, 0.15 + 3.2j], [120.0, 300.0]),
    ([1.6, 1.33, 1.0], [400.0]),
    ([1, 2.4, 1.46, 2.4, 1.46, 1.52], [57.0, 94.0, 57.0, 94.0]),
]


def f32c(z):
    z = complex(z)
    return complex(np.float32(z.real), np.float32(z.imag))


def literal(z):
    z = f32c(z)
    if z.imag == 0:
        return "%r" % z.real
    return "%r%s%ri" % (z.real, "+" if z.imag >= 0 else "-", abs(z.imag))


def stack_scene(ns, ds):
    text = ("Render 4 4 2 4 360 830\nMaterial f thinfilm [ %s ] [ %s ]\nObject o sphere f (0 0 0) 1\n"
            "Instance o default (0 0 0) (0 0 0) (1 1 1)\nBackground equirectangular 0\n"
            "Camera (0 0 -3) (-1 1 -2) (2 0 0) (0 -2 0)\n") % (" ".join(literal(z) for z in ns),
                                                              " ".join("%r" % d for d in ds))
    return lambda_amd.Scene.parse(text, "stack.lrd")


def stack_array(ns, ds):
    return np.array([len(ds)] + [v for z in ns for v in (f32c(z).real, f32c(z).imag)] + list(ds), np.float32)


def queries(n, seed):
    """Uniform cos(theta) plus a log-spaced grazing set, both orientations."""
    rng = np.random.default_rng(seed)
    q = np.zeros((n, 7), np.float32)
    q[:, 0] = np.concatenate([rng.random(n - 200), np.geomspace(1e-4, 3e-2, 200)])
    q[:, 1] = rng.uniform(360, 830, n)
    q[:, 2] = rng.integers(0, 2, n)
    return q


def within_tolerance(got, fp64, ambiguous):
    """Check (b): |R - R_fp64| <= 1e-5, relative above 1 (total internal
    reflection from the back makes Lambda's |M01/M00| exceed 1), on every
    query whose square-root branches are not decided by rounding noise."""
    ok = ~ambiguous
    err = np.abs(got.astype(np.float64) - fp64) / np.maximum(1.0, np.abs(fp64))
    return float(err[ok].max()), int(ambiguous.sum())
