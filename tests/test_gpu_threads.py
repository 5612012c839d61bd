"""Concurrent calls from several host threads, each on its own stream (run on
an MI355X: -m gpu).  The C ABI is reentrant: its only shared state is caches
of device facts filled on first use.  Every thread's results must equal the
same calls made afterwards from one thread, bit for bit (all products are
deterministic for a given shape and kernel)."""
import threading

import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
from tsm_benchmark_amd import _lib, ops  # noqa: E402

SHAPES = [(torch.float64, 8, 8), (torch.float32, 64, 64), (torch.bfloat16, 32, 16),
          (torch.float64, 64, 64), (torch.float32, 4, 32), (torch.bfloat16, 64, 64)]


def work(seed):
    """A^T B, A^T A, the backward pass, Y = A X and a batched launch over
    every shape; returns host copies of the results."""
    out, As, Bs = [], [], []
    for i, (dt, k, m) in enumerate(SHAPES):
        n = 20011 + 977 * i
        A = ops.random_matrix(n, k, dt, seed=seed, stream_id=1 + 2 * i)
        B = ops.random_matrix(n, m, dt, seed=seed, stream_id=2 + 2 * i)
        gdt = torch.float64 if dt == torch.float64 else torch.float32
        G = torch.linspace(-1.0, 1.0, k * m, dtype=gdt, device="cuda").reshape(k, m)
        out.append(ops.atb(A, B))
        out.append(ops.ata(A))
        out.extend(ops.atb_backward(A, B, G))
        out.append(ops.tsmm(A, G))
        if dt == torch.float64:  # one batch shares the input type
            As.append(A)
            Bs.append(B)
    out.extend(ops.atb_batched(As, Bs))
    return [x.cpu() for x in out]


def test_threads_on_own_streams_match_one_thread():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    torch.cuda.set_device(0)
    _lib.lib()
    seeds = [oracle.SEED + 101 * i for i in range(6)]
    results, errors = {}, []

    def run(seed):
        try:
            torch.cuda.set_device(0)
            stream = torch.cuda.Stream()
            with torch.cuda.stream(stream):
                r = work(seed)
            stream.synchronize()
            results[seed] = r
        except Exception as e:  # reported below, in the main thread
            errors.append((seed, repr(e)))

    threads = [threading.Thread(target=run, args=(s,)) for s in seeds]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=240)
    assert not any(t.is_alive() for t in threads), "a thread did not finish"
    assert not errors, errors
    for s in seeds:
        ref = work(s)
        got = results[s]
        assert len(got) == len(ref)
        for i, (x, y) in enumerate(zip(got, ref)):
            assert torch.equal(x, y), (s, i)
