"""Request server (dxmc_host_server_*): continuous batching over one GPU.

CPU: without a GPU the server cannot be created (no CPU fallback).
GPU: a request's result is the same alone and in a mix of concurrent
requests (identical histories: equal step/event counts and per-voxel event
counts; dose equal up to the order of FP64 atomic additions, 1e-10
relative); admitted requests overlap on the device; admission by memory
serialises requests under a small budget without changing results; a request
larger than the budget is refused.
"""
import time

import numpy as np
import pytest

from dxmclib_amd import backend, capi, workloads


def _request(side, kvp, histories, dx, seed, spacing=1.5):
    req = {"side": side, "spacing": spacing, "kvp": kvp, "histories": histories, "dx": dx, "seed": seed}
    mat, dens, names = workloads.cube_phantom(side, req["spacing"], seed)
    world = backend.host_world(mat, dens, (req["spacing"],) * 3, names)
    return world, workloads.request_spec(req)


def test_server_needs_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(backend.DxmcError):
        backend.Server()


def _run_all(server, reqs, seeds):
    ids = [server.submit(w, s, seed=sd) for (w, s), sd in zip(reqs, seeds)]
    out = []
    for i in ids:
        assert server.wait(i, 120.0) == capi.REQUEST_DONE
        out.append(server.collect(i))
    return out


@pytest.mark.gpu
def test_same_alone_and_in_mix():
    target = _request(40, 100.0, 400_000, True, 11)
    others = [_request(48, 120.0, 2_000_000, False, 12), _request(32, 80.0, 600_000, True, 13), _request(56, 140.0, 1_500_000, False, 14),
              _request(36, 90.0, 800_000, True, 15), _request(44, 110.0, 1_000_000, False, 16)]
    with backend.Server(slots=4, prep_threads=4) as srv:
        alone = _run_all(srv, [target], [7])[0]
        # hold admission until the whole mix is prepared, then admit it at once
        srv.pause(True)
        reqs = others[:3] + [target] + others[3:]
        ids = [srv.submit(w, s, seed=sd) for (w, s), sd in zip(reqs, [21, 22, 23, 7, 24, 25])]
        for i in ids:
            for _ in range(600):
                if srv.wait(i, 0.0) == capi.REQUEST_READY:
                    break
                time.sleep(0.05)
            assert srv.wait(i, 0.0) == capi.REQUEST_READY
        srv.pause(False)
        mix = []
        for i in ids:
            assert srv.wait(i, 120.0) == capi.REQUEST_DONE
            mix.append(srv.collect(i))
    got = mix[3]
    a, b = alone["info"], got["info"]
    print(f"alone: steps {a.steps} events {a.events}; in mix: steps {b.steps} events {b.events} (slot {b.slot})", flush=True)
    assert a.steps == b.steps and a.events == b.events and a.histories == b.histories
    assert np.array_equal(alone["n_events"], got["n_events"])
    assert np.allclose(alone["dose"], got["dose"], rtol=1e-10, atol=0)
    assert np.allclose(alone["variance"], got["variance"], rtol=1e-10, atol=0)
    # the admitted requests ran concurrently on the device
    spans = sorted((r["info"].admitted, r["info"].finished) for r in mix)
    overlaps = sum(1 for (s0, e0), (s1, e1) in zip(spans, spans[1:]) if s1 < e0)
    print("spans", [(round(s, 3), round(e, 3)) for s, e in spans], flush=True)
    assert overlaps >= 1
    assert len({r["info"].slot for r in mix}) >= 2


@pytest.mark.gpu
def test_admission_by_memory():
    reqs = [_request(40, 100.0, 300_000, True, 31), _request(40, 120.0, 300_000, False, 32), _request(40, 80.0, 300_000, True, 33)]
    with backend.Server(slots=4) as srv:
        free = _run_all(srv, reqs, [1, 2, 3])
    # the memory each request is charged; any two exceed the budget below
    one = max(r["info"].device_bytes for r in free)
    assert 2 * min(r["info"].device_bytes for r in free) > 1.5 * one
    with backend.Server(slots=4, memory_budget=int(one * 1.5)) as srv:
        small = _run_all(srv, reqs, [1, 2, 3])
        # never more than one admitted at a time under this budget
        spans = sorted((r["info"].admitted, r["info"].finished) for r in small)
        assert all(s1 >= e0 for (s0, e0), (s1, e1) in zip(spans, spans[1:]))
        big_world, big_spec = _request(128, 100.0, 1000, True, 34)
        with pytest.raises(backend.DxmcError):
            srv.submit(big_world, big_spec)
    for x, y in zip(small, free):
        assert x["info"].steps == y["info"].steps and x["info"].events == y["info"].events
        assert np.allclose(x["dose"], y["dose"], rtol=1e-10, atol=0)


@pytest.mark.gpu
def test_server_matches_transport_call():
    """A served request equals the same Transport call made directly."""
    world, spec = _request(32, 120.0, 200_000, False, 41)
    with backend.Server(slots=2) as srv:
        served = _run_all(srv, [(world, spec)], [5])[0]
    direct = backend.transport(world, spec, precision=capi.PRECISION_FP32, output_mode=capi.OUTPUT_EV_PER_HISTORY, seed=5)
    assert served["info"].steps == direct["stats"].steps and served["info"].events == direct["stats"].events
    assert np.allclose(served["dose"], direct["dose"], rtol=1e-10, atol=0)


@pytest.mark.gpu
def test_full_size_request_same_alone_and_in_mix():
    """The serving stream at its real size: a 256^3 CT-axial request with
    1e8 histories (the stream's largest side and history count), alone and
    admitted together with five requests of 128-224 voxels a side and 1e7
    to 5e7 histories. Same histories (step and event counts, per-voxel
    events) and dose up to the order of FP64 atomic additions."""
    print("full-size request alone", flush=True)
    target = _request(256, 120.0, 100_000_000, False, 51, spacing=1.2)
    others = [_request(224, 140.0, 50_000_000, True, 52, spacing=1.8), _request(128, 80.0, 10_000_000, False, 53, spacing=2.0),
              _request(192, 100.0, 30_000_000, True, 54, spacing=1.4), _request(160, 130.0, 20_000_000, False, 55, spacing=1.1),
              _request(200, 90.0, 40_000_000, True, 56, spacing=1.6)]
    with backend.Server(slots=4, prep_threads=8) as srv:
        alone = _run_all(srv, [target], [9])[0]
        print("full-size request in a mix", flush=True)
        mix = _run_all(srv, others[:2] + [target] + others[2:], [61, 62, 9, 63, 64, 65])
    got = mix[2]
    a, b = alone["info"], got["info"]
    print(f"alone: {a.histories} histories, steps {a.steps} events {a.events}; in mix: steps {b.steps} events {b.events}", flush=True)
    assert a.histories == b.histories >= 99_000_000
    assert a.steps == b.steps and a.events == b.events
    assert np.array_equal(alone["n_events"], got["n_events"])
    assert np.allclose(alone["dose"], got["dose"], rtol=1e-10, atol=0)
    assert alone["dose"].size == 256 ** 3 and alone["dose"].sum() > 0
