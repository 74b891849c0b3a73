"""lambda_hip_closest_hit_device, the closest-hit query that allocates
nothing (its traversal-stack overflow goes to a fixed per-device scratch
area), against lambda_hip_closest_hit on the same rays: two calls in a row,
the second smaller, on a scene whose traversal stack spills past LDS.  The
test allocates the device buffers itself through the HIP runtime."""
import ctypes as C
import os

import numpy as np
import pytest

import lambda_amd
from lambda_amd import _abi
from lambda_amd import scenes as gen

pytestmark = pytest.mark.gpu


def hip_runtime():
    for name in ("libamdhip64.so.7", "libamdhip64.so"):
        try:
            return C.CDLL(name, mode=C.RTLD_GLOBAL)
        except OSError:
            pass
    raise RuntimeError("HIP runtime not found")


class DeviceArray:
    def __init__(self, hip, nbytes):
        self.hip, self.nbytes, self.p = hip, nbytes, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), C.c_size_t(nbytes)) == 0

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0

    def download(self, a):
        assert self.hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(a.nbytes), 2) == 0
        return a

    def free(self):
        self.hip.hipFree(self.p)


@pytest.fixture(scope="module")
def mesh(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("noalloc"))
    lrd = gen.write_instanced_mesh(d)
    cwd = os.getcwd()
    os.chdir(d)
    try:
        s = lambda_amd.Scene.load(os.path.basename(lrd))
    finally:
        os.chdir(cwd)
    dev = lambda_amd.DeviceScene(s)
    yield dev
    dev.close()


def rays_in_box(dev, n, seed):
    info = dev.info()
    lo, hi = np.array(info["bounds"][:3]), np.array(info["bounds"][3:])
    rng = np.random.default_rng(seed)
    p = lo + (hi - lo) * rng.random((n, 3))
    q = lo + (hi - lo) * rng.random((n, 3))
    d = (q - p) / np.linalg.norm(q - p, axis=1, keepdims=True)
    return np.concatenate([p, d], 1).astype(np.float32)


def call(dev, d_rays, n, d_hits):
    err = C.create_string_buffer(512)
    rc = _abi.hip().lambda_hip_closest_hit_device(dev._h, d_rays, n, d_hits, None, err, 512)
    return rc, err.value.decode()


def test_no_alloc_closest_hit_matches_allocating_entry(mesh):
    info = mesh.info()
    print("stack need %d (LDS %d)" % (info["stack_need"], info["stack_lds"]), flush=True)
    assert info["stack_lds"] < info["stack_need"] <= 48
    hip = hip_runtime()
    n1, n2 = 150000, 3001
    d_rays = DeviceArray(hip, n1 * 24)
    d_hits = DeviceArray(hip, n1 * C.sizeof(_abi.Hit))
    try:
        for n, seed in ((n1, 31), (n2, 32)):
            rays = rays_in_box(mesh, n, seed)
            want = mesh.closest_hit(rays)
            d_rays.upload(rays)
            # sentinel beyond the rays of this call: must stay untouched
            sentinel = np.full(n1, -7, _abi.HIT_DTYPE)
            d_hits.upload(sentinel)
            rc, msg = call(mesh, d_rays.p, n, d_hits.p)
            assert rc == 0, msg
            assert hip.hipDeviceSynchronize() == 0
            got = d_hits.download(np.zeros(n1, _abi.HIT_DTYPE))
            print("n=%d: %d hits" % (n, int(want["hit"].sum())), flush=True)
            assert want["hit"].sum() > n // 10
            assert got[:n].tobytes() == want.tobytes()
            assert got[n:].tobytes() == sentinel[n:].tobytes()
        # a host pointer is refused before any launch
        host = np.zeros(12, np.float32)
        rc, msg = call(mesh, host.ctypes.data_as(C.c_void_p), 2, d_hits.p)
        assert rc == 1 and "device memory" in msg
        rc, msg = call(mesh, d_rays.p, -1, d_hits.p)
        assert rc == 1 and "invalid arguments" in msg
    finally:
        d_rays.free()
        d_hits.free()
