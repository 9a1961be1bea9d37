"""Serving engine on the GPU: continuous batching of 2-D jobs of uneven shapes
(elements, samples, grid, coherence factor, storage), admission by queue
length and by device-memory budget.  Every job's images in the mix must be
bit-identical to the same job run alone on a processor handle, and the alone
run is pinned to the oracle by the parity tests."""
import sys

import numpy as np
import pytest

from us_lab4a_amd import _lib, workloads
from us_lab4a_amd.serve import Engine, Job2D
from us_lab4a_amd.sta import CoherenceFactor, HipSTAProcessor

pytestmark = pytest.mark.gpu

MM = 1e-3
CFS = ["none", "phase_coherence_factor", "sign_coherence_factor"]


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def mix(n=8, seed=7):
    """Seeded jobs of uneven shapes (small enough for a test)."""
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n):
        N = int(rng.choice([16, 24, 32, 48]))
        ns = int(rng.choice([1024, 1536, 2048]))
        n_x, n_z = int(rng.choice([48, 64, 96])), int(rng.choice([64, 80, 128]))
        wl = workloads.Workload2D(name=f"mix{i}", num_elements=N, num_samples=ns, upsampling=int(rng.choice([2, 4])),
                                  cf=CFS[i % 3], grid_x=(-6 * MM, 6 * MM), grid_z=(5 * MM, 35 * MM), n_x=n_x,
                                  n_z=n_z, points=6, point_x=(-5 * MM, 5 * MM), point_z=(8 * MM, 30 * MM),
                                  seed_points=100 + i, seed_noise=200 + i)
        storage = "f16" if i % 2 else "f32"
        jobs.append((wl, storage, wl.echoes_cpu()))
    return jobs


def alone(wl, storage, echoes):
    p = HipSTAProcessor(wl.sta_config(), None, wl.upsampling, CoherenceFactor(wl.cf, wl.cf_param), wl.peak_offset,
                        wl.rx_apodization(), tx_apod=wl.tx_apodization(), storage=storage)
    p.set_grid(wl.grid())
    v, f = p.process_echoes(echoes)
    p.close()
    return v, f


def make_job(wl, storage, echoes):
    npix = wl.n_x * wl.n_z
    return Job2D(config=wl.sta_config(), upsampling=wl.upsampling, coherence=CoherenceFactor(wl.cf, wl.cf_param),
                 grid=wl.grid(), rx_apod=wl.rx_apodization(), echoes=echoes,
                 value=np.empty(npix, np.float32), factor=np.empty(npix, np.float32), storage=storage)


@pytest.fixture(scope="module")
def jobs_and_refs():
    jobs = mix()
    refs = [alone(*j) for j in jobs]
    return jobs, refs


def test_mix_matches_alone(jobs_and_refs):
    jobs, refs = jobs_and_refs
    say("serving mix of", len(jobs), "jobs, 4 in flight")
    with Engine(max_inflight=4) as eng:
        # every shape twice: the second pass reuses resident handles
        tickets = []
        for rep in range(2):
            for k, (wl, st, e) in enumerate(jobs):
                j = make_job(wl, st, e)
                tickets.append((k, j, eng.submit(j)))
        for k, j, t in tickets:
            s = eng.wait(t)
            assert s["ok"] and s["status"] == 0, s
            v_ref, f_ref = refs[k]
            np.testing.assert_array_equal(j.value.reshape(v_ref.shape), v_ref)
            np.testing.assert_array_equal(j.factor.reshape(f_ref.shape), f_ref)
        st = eng.stats()
        say("resident bytes", st["resident_bytes"])
        assert st["inflight"] == 0 and st["queued"] == 0


def test_device_resident_job(jobs_and_refs):
    import torch
    jobs, refs = jobs_and_refs
    wl, st, e = jobs[2]
    npix = wl.n_x * wl.n_z
    d_e = torch.from_numpy(e).cuda()
    d_v = torch.empty(npix, device="cuda")
    d_f = torch.empty(npix, device="cuda")
    with Engine(max_inflight=2) as eng:
        j = Job2D(config=wl.sta_config(), upsampling=wl.upsampling, coherence=CoherenceFactor(wl.cf, wl.cf_param),
                  grid=wl.grid(), rx_apod=wl.rx_apodization(), echoes=d_e, value=d_v, factor=d_f, storage=st)
        s = eng.wait(eng.submit(j))
        assert s["ok"], s
    np.testing.assert_array_equal(d_v.cpu().numpy().reshape(refs[2][0].shape), refs[2][0])
    np.testing.assert_array_equal(d_f.cpu().numpy().reshape(refs[2][1].shape), refs[2][1])


def test_admission_queue_full_is_busy(jobs_and_refs):
    jobs, _ = jobs_and_refs
    wl, st, e = jobs[0]
    busy = False
    held = []
    with Engine(max_inflight=1, max_queued=1) as eng:
        for _ in range(16):
            j = make_job(wl, st, e)
            try:
                held.append(eng.submit(j))
            except _lib.BusyError:
                busy = True
                break
        for t in held:
            assert eng.wait(t)["ok"]
    assert busy


def footprint_estimate(wl, storage):
    """The engine's admission estimate (serve.cpp estimate_bytes)."""
    ntx, nrx, L = wl.num_tx, wl.num_elements, wl.num_samples * wl.upsampling
    import oracle
    nf = max(oracle.next_fast_even(L + 1024), 2048)
    npix = wl.n_x * wl.n_z
    ntx_pad = 32 if ntx <= 32 else 64 if ntx <= 64 else 128
    return (ntx * nrx * L * (8 if storage == "f32" else 4) + nrx * npix * 4 + npix * 8
            + (npix // 256 + 1) * (nrx * ntx_pad * 8 + nrx * 8) + min(ntx * nrx, 1024) * nf * 28
            + ntx * nrx * wl.num_samples * 4 + npix * 8 + (16 << 20))


def test_memory_budget_evicts_and_refuses(jobs_and_refs):
    jobs, refs = jobs_and_refs
    # a budget that holds one handle at a time: shapes alternate, idle
    # handles are evicted, jobs wait for memory
    budget = int(1.05 * max(footprint_estimate(jobs[k][0], jobs[k][1]) for k in (0, 1, 3)))
    say("budget", budget)
    with Engine(max_inflight=2, hbm_budget=budget) as eng:
        ts = [(k, make_job(*jobs[k])) for k in (0, 1, 0, 3, 1)]
        ts = [(k, j, eng.submit(j)) for k, j in ts]
        for k, j, t in ts:
            s = eng.wait(t)
            assert s["ok"], s
            np.testing.assert_array_equal(j.value.reshape(refs[k][0].shape), refs[k][0])
    # a budget below any job's footprint: the job fails cleanly
    with Engine(max_inflight=1, hbm_budget=1 << 20) as eng:
        s = eng.wait(eng.submit(make_job(*jobs[0])))
        assert not s["ok"] and s["status"] == _lib.MISTA_E_NO_MEMORY
