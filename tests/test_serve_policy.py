"""Admission, co-batching and retirement of the request loop, on the CPU with
a stand-in backend (the GPU path is covered by tests/test_gpu_serve.py)."""
import numpy as np

from axisem_amd.serve import RequestLoop


class FakeSolver:
    """A request 'mesh=<m>;niter=<n>;wave=<bytes>;meshb=<bytes>'."""

    def __init__(self, spec, log):
        kv = dict(x.split("=") for x in spec.split(";"))
        self.mesh = kv["mesh"]
        self.dims = {"niter": int(kv["niter"]), "nel_solid": 4, "nel_fluid": 0,
                     "num_rec": 1}
        self.wave = int(kv["wave"])
        self.meshb = int(kv["meshb"])
        self.iter = 0
        self.log = log
        self.closed = False

    def close(self):
        self.closed = True


class FakeBatch:
    def __init__(self, sols, log):
        self.sols = sols
        log.append(("batch", tuple(s.mesh for s in sols), len(sols)))

    def step(self, n):
        out = []
        for s in self.sols:
            assert s.iter + n <= s.dims["niter"]
            out.append(np.arange(s.iter + 1, s.iter + n + 1, dtype=np.float32)
                       .reshape(n, 1, 1).repeat(3, axis=2))
            s.iter += n
        return out

    def close(self):
        pass


class FakeBackend:
    def __init__(self):
        self.log = []
        self.resident = {}

    def load(self, spec):
        return FakeSolver(spec, self.log)

    def key(self, sol):
        return (hash(sol.mesh) & 0xffffffff, 0)

    def batch_key(self, sol):
        # "<mesh>:<source type>": one device mesh part per type, one batch
        # for every type on a mesh
        return (hash(sol.mesh.split(":")[0]) & 0xffffffff, 1)

    def footprint(self, sol):
        return sol.meshb, sol.wave

    def mesh_get(self, sol):
        self.log.append(("mesh_get", sol.mesh))
        return sol.mesh

    def mesh_bind(self, sol, mesh):
        assert mesh == sol.mesh
        self.log.append(("bind", sol.mesh))

    def mesh_release(self, mesh):
        self.log.append(("release", mesh))

    def batch(self, sols, width):
        return FakeBatch(sols, self.log)


def _spec(mesh, niter, wave=10, meshb=100):
    return "mesh=%s;niter=%d;wave=%d;meshb=%d" % (mesh, niter, wave, meshb)


def test_every_request_gets_its_whole_seismogram():
    be = FakeBackend()
    loop = RequestLoop(hbm_budget_bytes=1000, chunk_steps=7, backend=be)
    reqs = [loop.submit(_spec("A", 30)), loop.submit(_spec("A", 45)),
            loop.submit(_spec("B", 12)), loop.submit(_spec("A", 20), arrival_s=0.0)]
    st = loop.run()
    assert st["completed"] == 4 and st["rejected"] == 0
    for r in reqs:
        np.testing.assert_array_equal(r.seis[:, 0, 0], np.arange(1, r.niter + 1))
    # the three requests on mesh A ran together, the mesh was built once
    assert st["max_cobatched"] == 3
    assert sum(1 for e in be.log if e == ("mesh_get", "A")) == 1
    assert loop.used == 0


def test_budget_limits_admission_and_idle_meshes_are_evicted():
    be = FakeBackend()
    # budget: one mesh (100) + two wavefields (2 x 10)
    loop = RequestLoop(hbm_budget_bytes=125, chunk_steps=5, backend=be)
    for m, n in (("A", 10), ("A", 10), ("A", 10), ("B", 10)):
        loop.submit(_spec(m, n))
    st = loop.run()
    assert st["completed"] == 4
    assert st["peak_device_bytes"] <= 125
    assert st["max_cobatched"] == 2
    # mesh A had to go before mesh B could come in
    rel = [i for i, e in enumerate(be.log) if e == ("release", "A")]
    get_b = [i for i, e in enumerate(be.log) if e == ("mesh_get", "B")]
    assert rel and get_b and rel[0] < get_b[0]


def test_a_request_that_cannot_fit_is_rejected():
    loop = RequestLoop(hbm_budget_bytes=50, backend=FakeBackend())
    r = loop.submit(_spec("A", 5, meshb=100))
    st = loop.run()
    assert st["rejected"] == 1 and r.error and r.seis is None


def test_max_members_per_batch():
    be = FakeBackend()
    loop = RequestLoop(hbm_budget_bytes=1e9, max_members=2, chunk_steps=4, backend=be)
    for k in range(5):
        loop.submit(_spec("A", 8 + k))
    st = loop.run()
    assert st["completed"] == 5 and st["max_cobatched"] == 2
    assert all(e[2] <= 2 for e in be.log if e[0] == "batch")


def test_source_types_of_one_mesh_share_a_batch():
    """The four runs of a moment tensor (different source types: separate
    device mesh parts) are stepped as one batch."""
    be = FakeBackend()
    loop = RequestLoop(hbm_budget_bytes=1000, chunk_steps=10, backend=be)
    for t in ("mono", "mono", "di", "quad"):
        loop.submit("mesh=A:%s;niter=20;wave=10;meshb=100" % t, 0.0)
    st = loop.run()
    assert st["completed"] == 4 and st["max_cobatched"] == 4
    assert ("batch", ("A:mono", "A:mono", "A:di", "A:quad"), 4) in be.log
    assert sorted(x[1] for x in be.log if x[0] == "mesh_get") == ["A:di", "A:mono", "A:quad"]
