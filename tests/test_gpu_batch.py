"""corrx_batch_* on the GPU: several sub-arrays (engines of different shapes and
forms) integrated and dumped in one run, on branch streams, directly or as a
replayed hipGraph. Every run's visibilities and integrators must equal, byte for
byte, what each engine gives alone (zero + corrx_integrate + corrx_dump on one
stream). The graph is replayed several times with the program's own eager work
and synchronizations between replays, and every replay is checked; new items
make a new capture.
"""

import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip(corr):
    return importlib.import_module("21cma_abf_correlator_amd.hip")


# (form, N, nf, chunk, T, rho)
SUBARRAYS = [(1, 5, 1024, 16, 40, 0.3), (2, 3, 2048, 8, 20, 0.1), (0, 4, 512, 32, 70, 0.5)]


class Sub:
    def __init__(self, corr, oracle, hip, i, form, N, nf, chunk, T, rho):
        self.x = oracle.gen_int8(100 + i, N, T, 2 * nf, rho, sigma=16.0, max_delay=4)
        self.N, self.nf, self.T, self.form = N, nf, T, form
        self.B = N * (N + 1) // 2
        self.pitch = T * 2 * nf
        self.eng = corr.Engine(N, nf, chunk, form)
        self.d_in = hip.Buffer.from_array(self.x)
        mp = (T + 7) // 8
        mask = np.full((N, mp), 0xFF, dtype=np.uint8)
        mask[0, 1] &= ~np.uint8(1 << 2)                 # station 0, frame 10 missing
        self.mask_pitch = mp
        self.d_mask = hip.Buffer.from_array(mask)
        if form != corr.FORM_FP32:
            self.eng.calibrate(self.d_in.ptr, self.pitch, T)
        self.nbytes = self.B * nf * self.eng.integrator_bytes()
        self.acc = hip.Buffer(self.nbytes)
        self.out = hip.Buffer(self.B * nf * 8)

    def item(self, corr, frames):
        return corr.Batch.item(self.d_in.ptr, self.pitch, frames, self.acc.ptr, d_out=self.out.ptr,
                               dump_frames=frames, d_mask=self.d_mask.ptr, mask_pitch=self.mask_pitch)

    def eager(self, hip, frames):
        self.acc.zero()
        self.eng.integrate(self.d_in.ptr, self.pitch, frames, self.acc.ptr, self.d_mask.ptr,
                           self.mask_pitch, 0)
        self.eng.dump(self.acc.ptr, frames, self.out.ptr)
        hip.synchronize()
        return self.result()

    def result(self):
        return (self.acc.to_numpy(np.uint8, (self.nbytes,)).copy(),
                self.out.to_numpy(np.uint8, (self.B * self.nf * 8,)).copy())

    def clear(self):
        self.acc.zero()
        self.out.zero()


def _same(a, b):
    return np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_batch_direct_and_graph(corr, oracle, hip):
    subs = [Sub(corr, oracle, hip, i, *s) for i, s in enumerate(SUBARRAYS)]
    frames = [s.T for s in subs]
    ref = [s.eager(hip, f) for s, f in zip(subs, frames)]
    batch = corr.Batch([s.eng for s in subs])
    stream = hip.Stream()
    items = [s.item(corr, f) for s, f in zip(subs, frames)]

    print("batch direct", flush=True)
    for s in subs:
        s.clear()
    batch.run(items, graph=False, stream=stream.handle)
    stream.synchronize()
    assert all(_same(s.result(), r) for s, r in zip(subs, ref))

    for rep in range(4):
        print(f"batch graph replay {rep}", flush=True)
        for s in subs:
            s.clear()
        # the program's own eager work between replays
        subs[0].eng.fengine(subs[0].d_in.ptr, subs[0].pitch, 4)
        hip.synchronize()
        batch.run(items, graph=True, stream=stream.handle)
        stream.synchronize()
        got = [s.result() for s in subs]
        assert all(_same(g, r) for g, r in zip(got, ref)), f"replay {rep} differs from the eager run"

    # new items (fewer frames in one sub-array): a new capture, again exact
    frames2 = [frames[0], frames[1] - 7, frames[2]]
    ref2 = [s.eager(hip, f) for s, f in zip(subs, frames2)]
    items2 = [s.item(corr, f) for s, f in zip(subs, frames2)]
    for rep in range(2):
        print(f"batch graph, new items, replay {rep}", flush=True)
        for s in subs:
            s.clear()
        batch.run(items2, graph=True, stream=stream.handle)
        stream.synchronize()
        assert all(_same(s.result(), r) for s, r in zip(subs, ref2))

    # and the visibilities themselves against the FP64 oracle
    for s, r in zip(subs, ref):
        valid = np.ones((s.N, s.T), dtype=np.uint8)
        valid[0, 10] = 0
        V = r[1].view(np.complex64).reshape(s.B, s.nf)
        ref64 = oracle.fx_f64(s.x, valid=valid)
        err = float(np.sqrt(np.sum(np.abs(V - ref64) ** 2) / np.sum(np.abs(ref64) ** 2)))
        print(f"  form {s.form} N={s.N} nf={s.nf}: vs FP64 {err:.3e}", flush=True)
        assert err < {0: 1e-5, 1: 1e-3, 2: 0.1}[s.form]
    batch.close()
    stream.destroy()
    for s in subs:
        s.eng.close()


def test_batch_rejections(corr, hip):
    a = corr.Engine(2, 512, 8, corr.FORM_FP16)
    b = corr.Engine(3, 512, 8, corr.FORM_FP32)
    with pytest.raises(corr.CorrError):
        corr.Batch([a, a])                                   # the same engine twice
    batch = corr.Batch([a, b])
    acc = hip.Buffer(6 * 512 * 16)
    d_in = hip.Buffer(3 * 8 * 1024)
    good = corr.Batch.item(d_in.ptr, 8 * 1024, 8, acc.ptr)
    bad_flags = corr.Batch.item(d_in.ptr, 8 * 1024, 8, acc.ptr)
    bad_flags.flags = 8
    no_out = corr.Batch.item(d_in.ptr, 8 * 1024, 8, acc.ptr)
    no_out.flags |= corr.BATCH_DUMP
    short = corr.Batch.item(d_in.ptr, 8 * 1024, 9, acc.ptr)  # pitch holds 8 frames
    for bad in (bad_flags, no_out, short):
        for graph in (False, True):
            with pytest.raises(corr.CorrError):
                batch.run([good, bad], graph=graph)
    batch.close()
    a.close()
    b.close()
