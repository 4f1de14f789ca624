"""Serving engine (continuous batching) through the C ABI.

* a job's field equals the one-shot ohx_pc_diffract result to FP64 rounding
  of the per-slice partial sums (rel 1e-12);
* the same job gives a bit-identical field alone and in a mix of other jobs
  (RS / Fresnel / matrix-core Fresnel, several resolutions) -- slices depend
  only on the job, each (job, tile) item has one owner;
* graph replay and direct launches give identical fields;
* admission by HBM budget, rejection of jobs that can never fit, and
  cancellation of queued jobs are forced from the host.
"""
import numpy as np
import pytest

from openholo_amd import capi, synth

pytestmark = pytest.mark.gpu
FP32 = capi.MODE_GPU | capi.MODE_FLOAT
MFMA3 = FP32 | capi.MODE_MFMA_BF16X3


def cloud(seed, n):
    return synth.random_cloud(seed, n, 4e-3, (0.49, 0.51))


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def oneshot(v, p, diff, mode):
    with capi.Context(0) as ctx:
        pn = (p.pn_x, p.pn_y)
        return capi.generate(ctx, v, pn, (p.pp_x, p.pp_y), [p.lambda_], diff, mode)[0].ravel()


JOBS = [
    # seed, points, resolution, pitch, lambda, diff, mode
    (1, 3000, (1920, 1080), (8e-6, 8e-6), 638e-9, 0, FP32),
    (2, 20000, (640, 480), (8e-6, 8e-6), 532e-9, 1, FP32),
    (3, 9000, (1920, 1080), (8e-6, 8e-6), 473e-9, 1, MFMA3),
    (4, 1500, (3840, 2160), (3.74e-6, 3.74e-6), 532e-9, 0, FP32),
    (5, 7000, (800, 600), (8e-6, 8e-6), 638e-9, 1, MFMA3),
]


def submit(srv, spec, want_u8=False):
    seed, n, pn, pp, lam, diff, mode = spec
    return srv.submit(cloud(seed, n), capi.params(pn, pp, lam), diff, mode, want_field=True, want_u8=want_u8)


def test_single_job_matches_oneshot():
    with capi.Server(slice_pairs=2e9, slice_pairs_mfma=5e9) as srv:
        for spec in JOBS[:3]:
            jid = submit(srv, spec)
            st, stats, f, _ = srv.wait(jid, 120)
            assert st == capi.OK
            assert stats.iterations >= 2          # really sliced
            seed, n, pn, pp, lam, diff, mode = spec
            ref = oneshot(cloud(seed, n), capi.params(pn, pp, lam), diff, mode)
            # direct: FP64 rounding of per-slice partial sums; matrix cores:
            # the FP32 split-K partials differ with the slicing (3-term error)
            tol = 1e-12 if mode == FP32 else 2e-5
            assert rel_l2(f, ref) < tol


@pytest.mark.parametrize("use_graph", [True, False])
def test_alone_equals_in_mix(use_graph):
    with capi.Server(use_graph=use_graph, slice_pairs=2e9, slice_pairs_mfma=5e9) as srv:
        alone = {}
        for k, spec in enumerate(JOBS):
            st, _, f, _ = srv.wait(submit(srv, spec), 120)
            assert st == capi.OK
            alone[k] = f
        ids = {k: submit(srv, spec) for k, spec in enumerate(JOBS)}      # all in flight together
        ids2 = {k: submit(srv, spec) for k, spec in reversed(list(enumerate(JOBS)))}
        for d in (ids, ids2):
            for k, jid in d.items():
                st, stats, f, _ = srv.wait(jid, 300)
                assert st == capi.OK
                assert np.array_equal(f, alone[k]), k
        c = srv.counters()
        assert c.jobs_done == 3 * len(JOBS)
        assert (c.graph_replays > 0) == use_graph


def test_fp64_jobs_in_mix():
    """MODE_DOUBLE jobs (the FP64 tiles of the multi-job kernel, 128x16) next to
    FP32 and matrix-core jobs: each equals the one-shot FP64 launch to FP64
    rounding and is bit-identical alone or in the mix."""
    fp64 = capi.MODE_GPU
    specs = [(61, 1800, (1920, 1080), (8e-6, 8e-6), 532e-9, 0, fp64),
             (62, 6000, (1000, 700), (8e-6, 8e-6), 638e-9, 1, fp64)]
    with capi.Server(slice_pairs=2e9, slice_pairs_mfma=5e9) as srv:
        alone = []
        for spec in specs:
            st, stats, f, _ = srv.wait(submit(srv, spec), 120)
            assert st == capi.OK and stats.iterations >= 2
            seed, n, pn, pp, lam, diff, mode = spec
            ref = oneshot(cloud(seed, n), capi.params(pn, pp, lam), diff, mode)
            assert rel_l2(f, ref) < 1e-12
            alone.append(f)
        ids = [submit(srv, s) for s in [JOBS[0], specs[0], JOBS[2], specs[1]]]
        res = [srv.wait(j, 300) for j in ids]
        assert all(r[0] == capi.OK for r in res)
        assert np.array_equal(res[1][2], alone[0]) and np.array_equal(res[3][2], alone[1])


def test_graph_and_direct_launch_agree():
    fields = []
    for g in (True, False):
        with capi.Server(use_graph=g, slice_pairs=2e9) as srv:
            st, _, f, _ = srv.wait(submit(srv, JOBS[0]), 120)
            assert st == capi.OK
            fields.append(f)
    assert np.array_equal(fields[0], fields[1])


def test_u8_output_matches_oracle(orc):
    with capi.Server(slice_pairs=2e9) as srv:
        st, _, f, u = srv.wait(submit(srv, JOBS[1], want_u8=True), 120)
        assert st == capi.OK
    want = orc.normalize([orc.encode(f, orc.ENCODE_PHASE)])[0]
    assert np.array_equal(u, want)


def test_admission_by_memory_budget():
    spec = JOBS[3]                       # 3840x2160: ~0.2 GB of field + buffers
    with capi.Server(mem_budget=64 << 20) as srv:
        with pytest.raises(capi.OhxError) as e:
            submit(srv, spec)
        assert e.value.status == capi.ERR_ADMISSION
    one_job = 8294400 * 25 + (64 << 20)
    with capi.Server(mem_budget=one_job, slice_pairs=2e9) as srv:
        ids = [submit(srv, spec) for _ in range(3)]
        res = [srv.wait(j, 300) for j in ids]
        assert all(r[0] == capi.OK for r in res)
        for a, b in zip(res, res[1:]):   # strictly one at a time
            assert b[1].admit_ms >= a[1].end_ms - 1e-6
        c = srv.counters()
        assert c.peak_reserved_bytes <= one_job and c.reserved_bytes == 0


def test_cancel_pending_job():
    spec = JOBS[3]
    one_job = 8294400 * 25 + (64 << 20)
    with capi.Server(mem_budget=one_job, slice_pairs=2e9) as srv:
        a = submit(srv, spec)
        b = submit(srv, spec)             # queued behind a (budget)
        assert srv.cancel(b) == capi.OK
        assert srv.wait(b, 60)[0] == capi.ERR_CANCELLED
        assert srv.wait(a, 300)[0] == capi.OK
        assert srv.cancel(12345) == capi.ERR_NOT_FOUND


def test_rs_matrix_core_jobs():
    """RS jobs with MODE_MFMA_BF16X3: the setup partitions the points on the
    device (points with partial masks -- here the near ones -- to the direct
    kernel); GEMM slices cover the others, then direct slices the rest.  The
    field is within 2e-5 of the one-shot launch (which splits the same points
    into different K slices) and bit-identical alone and in a mix."""
    near = synth.random_cloud(71, 500, 2e-3, (2e-3, 4e-3))
    far = synth.random_cloud(72, 6000, 4e-3, (0.49, 0.51))
    v = np.concatenate([far[:3000], near, far[3000:]])
    pn, pp, lam = (1920, 1080), (8e-6, 8e-6), 532e-9
    p = capi.params(pn, pp, lam)
    with capi.Server(slice_pairs=2e9, slice_pairs_mfma=5e9) as srv:
        st, stats, alone, _ = srv.wait(srv.submit(v, p, 0, MFMA3, want_field=True), 120)
        assert st == capi.OK and stats.iterations >= 3
        ref = oneshot(v, p, 0, MFMA3)
        assert rel_l2(alone, ref) < 2e-5
        ids = [submit(srv, JOBS[0]), srv.submit(v, p, 0, MFMA3, want_field=True), submit(srv, JOBS[2])]
        res = [srv.wait(j, 300) for j in ids]
        assert all(r[0] == capi.OK for r in res)
        assert np.array_equal(res[1][2], alone)


def test_rs_matrix_core_job_without_matrix_core_points():
    """An RS job with MODE_MFMA_BF16X3 none of whose points qualify (all near,
    partial masks): it runs as direct slices only and equals the FP32 direct
    job to FP64 rounding of the slice partial sums."""
    v = synth.random_cloud(73, 800, 2e-3, (2e-3, 4e-3))
    pn, pp, lam = (640, 480), (8e-6, 8e-6), 473e-9
    p = capi.params(pn, pp, lam)
    with capi.Server(slice_pairs=2e8, slice_pairs_mfma=5e9) as srv:
        st, stats, f_mfma, _ = srv.wait(srv.submit(v, p, 0, MFMA3, want_field=True), 120)
        assert st == capi.OK and stats.iterations >= 2
        st, _, f_direct, _ = srv.wait(srv.submit(v, p, 0, FP32, want_field=True), 120)
        assert st == capi.OK
    assert rel_l2(f_mfma, f_direct) < 1e-12
