"""tm_peak_caller_fold_noalloc (the peak-caller step on the library's static
per-GPU scratch, for callers that cannot allocate) against
tm_peak_caller_fold with a caller workspace, through the C-ABI: identical
outputs bit for bit on two calls in a row, the second with fewer candidates
and a non-empty state; steps beyond the scratch are refused before any
launch."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _volume(t, shape, seed):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(shape).astype(np.float32) * 0.1
    for _ in range(40):  # planted peaks
        p = tuple(int(rng.integers(2, n - 2)) for n in shape)
        v[p] = 1.0 + rng.random()
    return t.tensor(v, device="cuda")


def _candidates(be, scores):
    return be.to_backend_array(be.max_filter_coordinates(scores, 3)).to(dtype=be._torch.int64).contiguous()


def _fold(lib_call, t, be, noalloc, cand, scores, state, rotation, num_peaks, min_distance, mbd, lo):
    from pytme_amd import _lib

    pos, rot, sc, det = state
    n_state = int(pos.shape[0])
    dev = scores.device
    out = (t.full((num_peaks, 3), -7, dtype=t.int64, device=dev), t.full((num_peaks, 3, 3), -7.0, device=dev),
           t.full((num_peaks,), -7.0, device=dev), t.full((num_peaks,), -7.0, device=dev))
    counts = t.full((3,), -7, dtype=t.int64, device=dev)
    ptr = lambda x: x.data_ptr() if x.numel() else None  # noqa: E731
    args = [cand.data_ptr(), int(cand.shape[0]), scores.data_ptr(), _lib.i64_array(scores.shape), ptr(pos), ptr(rot),
            ptr(sc), ptr(det), n_state, rotation.data_ptr(), num_peaks, min_distance, mbd, 1, float(lo), 0, 0.0,
            out[0].data_ptr(), out[1].data_ptr(), out[2].data_ptr(), out[3].data_ptr(), counts.data_ptr()]
    stream = be._stream()
    if noalloc:
        lib_call("tm_peak_caller_fold_noalloc", *args, stream)
    else:
        nbytes = int(_lib.lib().tm_peak_caller_fold_workspace_bytes(int(cand.shape[0]), n_state, num_peaks))
        ws = t.full((nbytes,), 0xAB, dtype=t.uint8, device=dev)
        lib_call("tm_peak_caller_fold", *args, ws.data_ptr(), nbytes, stream)
    t.cuda.synchronize()
    return out, counts


def test_noalloc_fold_matches_workspace_fold():
    import torch

    from pytme_amd import _lib
    from pytme_amd.backends import backend as be

    be.change_backend("hip", device="cuda")
    R = torch.tensor(np.eye(3, dtype=np.float32), device="cuda")
    empty = (torch.zeros((0, 3), dtype=torch.int64, device="cuda"), torch.zeros((0, 3, 3), device="cuda"),
             torch.zeros(0, device="cuda"), torch.zeros(0, device="cuda"))
    # call 1: 96x80x64 volume, empty state; call 2: a smaller volume and
    # the first call's rows as the state
    s1 = _volume(torch, (96, 80, 64), 1)
    c1 = _candidates(be, s1)
    print(f"no-alloc fold call 1: {int(c1.shape[0])} candidates", flush=True)
    a_out, a_cnt = _fold(_lib.call, torch, be, False, c1, s1, empty, R, 64, 4, 2, 0.05)
    b_out, b_cnt = _fold(_lib.call, torch, be, True, c1, s1, empty, R, 64, 4, 2, 0.05)
    assert torch.equal(a_cnt, b_cnt) and int(a_cnt[2]) > 0
    for a, b in zip(a_out, b_out):
        assert torch.equal(a, b)
    n_out = int(a_cnt[2])
    state = tuple(x[:n_out].contiguous() for x in a_out)

    s2 = _volume(torch, (40, 36, 32), 2)
    c2 = _candidates(be, s2)
    assert c2.shape[0] < c1.shape[0]
    R2 = torch.tensor(np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]], dtype=np.float32), device="cuda")
    print(f"no-alloc fold call 2: {int(c2.shape[0])} candidates, {n_out} state rows", flush=True)
    a_out, a_cnt = _fold(_lib.call, torch, be, False, c2, s2, state, R2, 32, 3, 1, 0.05)
    b_out, b_cnt = _fold(_lib.call, torch, be, True, c2, s2, state, R2, 32, 3, 1, 0.05)
    assert torch.equal(a_cnt, b_cnt) and int(a_cnt[2]) > 0
    for a, b in zip(a_out, b_out):
        assert torch.equal(a, b)


def test_noalloc_fold_refuses_steps_beyond_the_scratch():
    import torch

    from pytme_amd import _lib
    from pytme_amd.backends import backend as be

    be.change_backend("hip", device="cuda")
    lib = _lib.lib()
    cap = int(lib.tm_peak_caller_fold_noalloc_max_bytes())
    n = 1
    while lib.tm_peak_caller_fold_workspace_bytes(n, 0, 1000) <= cap:
        n *= 2
    assert lib.tm_peak_caller_fold_workspace_bytes(n, 0, 1000) > cap
    # the size check comes before any pointer is read or kernel launched
    dummy = torch.zeros(16, dtype=torch.int64, device="cuda")
    scores = torch.zeros((4, 4, 4), device="cuda")
    out = torch.zeros(64, device="cuda")
    with pytest.raises(ValueError, match="scratch"):
        _lib.call("tm_peak_caller_fold_noalloc", dummy.data_ptr(), n, scores.data_ptr(), _lib.i64_array((4, 4, 4)),
                  None, None, None, None, 0, out.data_ptr(), 1000, 3, 0, 0, 0.0, 0, 0.0, dummy.data_ptr(),
                  out.data_ptr(), out.data_ptr(), out.data_ptr(), dummy.data_ptr(), be._stream())
