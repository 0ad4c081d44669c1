"""Scratch of the stream-ordered entry points: one buffer per (device,
stream).  Launches in flight on different streams (and from different host
threads) must not see each other's partial sums, a buffer that grows must
not be freed under a launch, nb_scratch_release frees everything, and a call
that would grow a buffer while its stream is captured is refused."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def disk(nb, torch, gpu, n, seed):
    pos, vel = nb.disk_ic(n, seed=seed)
    P = torch.from_numpy(np.stack(pos)).to(gpu)
    V = torch.from_numpy(np.stack(vel)).to(gpu)
    return P, V


def cotangents(torch, gpu, n, seed):
    rng = np.random.default_rng(seed)
    return [torch.from_numpy(rng.normal(size=(3, n)).astype(np.float32)).to(gpu)
            for _ in range(2)]


def fwd(nb, torch, params, P, V, stream):
    v = V.clone()
    out = torch.empty_like(P)
    nb.forward_step(params, list(P), list(v), list(out), stream=stream)
    return out, v


def adj(nb, torch, params, P, lxn, lvn, stream):
    lx, lv = torch.empty_like(lxn), torch.empty_like(lvn)
    nb.adjoint_step(params, P, lxn, lvn, lx, lv, stream=stream)
    return lx, lv


def same(a, b):
    return all(torch_equal(x, y) for x, y in zip(a, b))


def torch_equal(x, y):
    import torch

    return bool(torch.equal(x, y))


def test_two_streams_in_flight(nb, gpu):
    """A forward step (symmetric shape, partial sums) on one stream and an
    adjoint step (symmetric shape, partial sums) on another, enqueued
    back to back several times so that they run together: every result
    equals the same call run alone."""
    import torch

    n1, n2 = 131072, 98304  # different sizes: different partial layouts
    p1 = nb.default_params(num_particles=n1)
    p2 = nb.default_params(num_particles=n2)
    P1, V1 = disk(nb, torch, gpu, n1, 11)
    P2, _ = disk(nb, torch, gpu, n2, 12)
    lxn, lvn = cotangents(torch, gpu, n2, 13)
    s0 = torch.cuda.current_stream()
    ref_f = fwd(nb, torch, p1, P1, V1, s0)
    ref_a = adj(nb, torch, p2, P2, lxn, lvn, s0)
    torch.cuda.synchronize()
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    outs_f, outs_a = [], []
    torch.cuda.synchronize()
    for _ in range(4):
        with torch.cuda.stream(sa):
            outs_f.append(fwd(nb, torch, p1, P1, V1, sa))
        with torch.cuda.stream(sb):
            outs_a.append(adj(nb, torch, p2, P2, lxn, lvn, sb))
    torch.cuda.synchronize()
    for k, (f, a) in enumerate(zip(outs_f, outs_a)):
        assert same(f, ref_f), f"forward on stream A, round {k}"
        assert same(a, ref_a), f"adjoint on stream B, round {k}"


def test_host_threads_on_own_streams(nb, gpu):
    """Two host threads, each on its own stream, each repeating a step of a
    different size: the results equal the serial ones."""
    import torch

    sizes = (65536, 81920)
    probs = []
    for k, n in enumerate(sizes):
        p = nb.default_params(num_particles=n)
        P, V = disk(nb, torch, gpu, n, 20 + k)
        probs.append((p, P, V))
    s0 = torch.cuda.current_stream()
    refs = [fwd(nb, torch, p, P, V, s0) for (p, P, V) in probs]
    torch.cuda.synchronize()
    results = [[], []]
    errors = []

    def work(k):
        try:
            s = torch.cuda.Stream()
            p, P, V = probs[k]
            with torch.cuda.stream(s):
                for _ in range(6):
                    results[k].append(fwd(nb, torch, p, P, V, s))
            s.synchronize()
        except Exception as e:  # reported by the main thread
            errors.append(e)

    threads = [threading.Thread(target=work, args=(k,)) for k in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    torch.cuda.synchronize()
    for k in range(2):
        assert len(results[k]) == 6
        for r in results[k]:
            assert same(r, refs[k]), sizes[k]


def test_release_frees_and_reallocates(nb, gpu):
    import torch

    n = 65536
    p = nb.default_params(num_particles=n)
    P, V = disk(nb, torch, gpu, n, 31)
    s = torch.cuda.current_stream()
    first = fwd(nb, torch, p, P, V, s)
    torch.cuda.synchronize()
    assert nb.scratch_held() > 0
    nb.scratch_release()
    assert nb.scratch_held() == 0
    again = fwd(nb, torch, p, P, V, s)
    torch.cuda.synchronize()
    assert same(first, again)
    assert nb.scratch_held() > 0


def test_growth_refused_under_capture(nb, gpu):
    """Growing a stream's buffer inside graph capture is refused; after one
    call of the size on that stream the captured call replays bit-equal to
    the direct call, replay after replay."""
    import torch

    n = 32768
    p = nb.default_params(num_particles=n)
    P, V = disk(nb, torch, gpu, n, 41)
    ref = fwd(nb, torch, p, P, V, torch.cuda.current_stream())
    torch.cuda.synchronize()
    nb.scratch_release()
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    v = V.clone()
    out = torch.empty_like(P)
    with pytest.raises(nb.NbError, match="before capturing"):
        with torch.cuda.graph(g, stream=s):
            nb.forward_step(p, list(P), list(v), list(out), stream=s)
    torch.cuda.synchronize()
    # set the stream's buffer up outside capture, then capture
    with torch.cuda.stream(s):
        fwd(nb, torch, p, P, V, s)
    s.synchronize()
    g = torch.cuda.CUDAGraph()
    v = V.clone()
    with torch.cuda.graph(g, stream=s):
        nb.forward_step(p, list(P), list(v), list(out), stream=s)
    for rep in range(3):
        v.copy_(V)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert same((out, v), ref), rep
