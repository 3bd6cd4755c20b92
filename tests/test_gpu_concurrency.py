"""Concurrent callers: the C-ABI is stream-ordered and its process-wide tables
(tuned variants, launch shapes per size class) sit behind a mutex, so host
threads driving their own streams must get exactly the bits a single thread
gets -- including when both make the first call of a size class at once, so
that both time the launch shapes on their own buffers (every shape gives the
same bits).  Two threads, each with its own stream, run the forward, the
8-bit forward and the adjoint on their own volumes, several times each."""
import threading

import numpy as np
import pytest
import torch

import gnlm

pytestmark = pytest.mark.gpu


def _work(dev, seed):
    rng = np.random.default_rng(seed)
    u = torch.from_numpy(rng.random((40, 44, 36), dtype=np.float32)).to(dev)
    u8 = torch.from_numpy(rng.integers(0, 256, (24, 30, 28), dtype=np.uint8)).to(dev)
    gy = torch.from_numpy(rng.standard_normal((40, 44, 36)).astype(np.float32)).to(dev)
    return u, u8, gy


def _run(u, u8, gy):
    y, Z = gnlm.nlm3d(u, 2, 3, 0.2, want_Z=True)
    y1 = gnlm.nlm3d(u, 1, 4, 0.15)
    y8 = gnlm.nlm3d_u8(u8, 8, 2, 25.0)
    gu, gh = gnlm.nlm3d_backward(u, y, Z, gy, 2, 3, 0.2)
    return [t.clone() for t in (y, Z, y1, y8, gu, gh)]


def test_two_threads_two_streams_same_bits(dev):
    works = [_work(dev, s) for s in (1, 2)]
    # the single-threaded reference, after a reset so that the threads below
    # also race on the first call of each size class
    gnlm.reset_tuning()
    ref = [_run(*w) for w in works]
    torch.cuda.synchronize()
    gnlm.reset_tuning()
    results = [[], []]
    errors = []

    def worker(i):
        try:
            s = torch.cuda.Stream(device=dev)
            with torch.cuda.stream(s):
                for _ in range(3):
                    results[i].append(_run(*works[i]))
            s.synchronize()
        except Exception as e:  # reported by the main thread
            errors.append(repr(e))

    print("case two threads, two streams", flush=True)
    threads = [threading.Thread(target=worker, args=(i,)) for i in (0, 1)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads), "worker did not finish"
    torch.cuda.synchronize()
    assert not errors, errors
    for i in (0, 1):
        assert len(results[i]) == 3
        for rep in results[i]:
            for got, want in zip(rep, ref[i]):
                assert torch.equal(got, want)
    gnlm.reset_tuning()
